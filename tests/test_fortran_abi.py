"""The ISO_C_BINDING Fortran binding on the CPU: bind(C) type sizes equal the C
structs, and the host-only plan returns the reference's golden n(Q) through it."""
import json
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "trj_analysis_amd", "fortran", "abi_check")


def test_fortran_abi_check():
    assert os.path.exists(EXE), "run __graft_entry__.build() first (amdflang)"
    out = subprocess.run([EXE], capture_output=True, text=True, check=True).stdout.split("\n")
    vals = {line.split()[0]: line.split()[1:] for line in out if line.strip()}
    assert vals["sizes"] == ["40", "128", "104"]
    assert vals["abi"] == ["1"]
    assert vals["plan"] == ["67", "14", "14"]
    with open(os.path.join(ROOT, "tests", "golden", "nq_golden.json")) as f:
        lj = [c for c in json.load(f)["cases"] if c["name"] == "3D/LJmix"][0]
    assert [int(v) for v in vals["nq"]] == lj["nq"][:16]
    assert int(vals["rdf"][0]) - 2 == lj["gmix_rows"]
