"""Extract golden vectors from the reference's shipped example outputs.

Also copies each case's input.nml (data) to inputs/<case>.nml.

Reads (as data, never as code) the n(Q) column of sq.dat and the row count of
gmixsim.dat in examples/ of the reference checkout, plus the /INPUT_RDF/ and
/INPUT_SQ/ values of the matching input.nml, and writes nq_golden.json next to
this script.  The integer n(Q) column is the only part of those outputs that is
reproducible without the (unshipped) trajectories: it depends on the box only
through dq = 2*pi/L, which the first Q column prints.

Usage: python tests/golden/make_golden.py [/path/to/reference]
"""
import json
import os
import re
import sys

CASES = [
    # name, directory, ndim
    ("3D/LJmix", "examples/3D/LJmix", 3),
    ("3D/salr", "examples/3D/salr", 3),
    ("3D/meam-dyn", "examples/3D/meam-dyn", 3),
    ("2D/salr", "examples/2D/salr", 2),
]


def _nml_value(text, group, key):
    m = re.search(r"&%s\b(.*?)\n\s*/" % group, text, re.S | re.I)
    if not m:
        return None
    for line in m.group(1).splitlines():
        line = line.split("!")[0]
        mm = re.match(r"\s*%s\s*=\s*(.+)$" % key, line, re.I)
        if mm:
            return float(mm.group(1).split()[0])
    return None


def main(ref):
    out = []
    for name, rel, ndim in CASES:
        d = os.path.join(ref, rel)
        with open(os.path.join(d, "input.nml")) as f:
            nml = f.read()
        q, nq = [], []
        with open(os.path.join(d, "sq.dat")) as f:
            for line in f:
                if line.startswith("#") or not line.strip():
                    continue
                cols = line.split()
                q.append(float(cols[0]))
                nq.append(int(cols[-1]))
        grows = 0
        gpath = os.path.join(d, "gmixsim.dat")
        if os.path.exists(gpath):
            with open(gpath) as f:
                grows = sum(1 for line in f if line.strip() and not line.startswith("#"))
        out.append({
            "name": name,
            "ndim": ndim,
            "dq_printed": q[0],
            "qmin": _nml_value(nml, "INPUT_SQ", "qmin"),
            "qmax": _nml_value(nml, "INPUT_SQ", "qmax"),
            "deltar": _nml_value(nml, "INPUT_RDF", "deltar"),
            "rcrdf": _nml_value(nml, "INPUT_RDF", "rcrdf"),
            "nq": nq,
            "gmix_rows": grows,
        })
        # the namelist itself, as a data fixture for the namelist reader
        nml_dir = os.path.join(os.path.dirname(os.path.abspath(__file__)), "inputs")
        os.makedirs(nml_dir, exist_ok=True)
        with open(os.path.join(nml_dir, name.replace("/", "_") + ".nml"), "w") as f:
            f.write(nml)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "nq_golden.json")
    with open(path, "w") as f:
        json.dump({"source": "reference examples/*/sq.dat n(Q) column, gmixsim.dat rows",
                   "cases": out}, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
