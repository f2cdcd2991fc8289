"""The oracle pins catch plausible mistakes: each case plants one mistake in a
copy of oracle/trj_oracle.c (compiled into a temporary directory, on the
CPU) and runs tests/test_oracle_pins.py against it, which must fail.  The
mistakes: N for N_j or N_1 in the partial columns, a dropped nq, Nconf or
<b>^2 in S_NN, swapped x1/x2 in S_cc, a dropped zgrid/Lz, pair factor or
Nconf in the slice g_xy, a wrong |q|^2 or shell rounding in the member rule."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "trj_oracle.c")

MUTATIONS = {
    "partial_N_for_Nj": ("mrow[1 + s] = sqfp[(int64_t)s * ns + i] / ((double)ntype[s]",
                         "mrow[1 + s] = sqfp[(int64_t)s * ns + i] / ((double)natoms"),
    "snn_without_nq": ("double snn = (sqf[i] / (baver * baver)) / ((double)natoms * (double)nconf * nqi);",
                       "double snn = (sqf[i] / (baver * baver)) / ((double)natoms * (double)nconf);"),
    "snn_without_nconf": ("double snn = (sqf[i] / (baver * baver)) / ((double)natoms * (double)nconf * nqi);",
                          "double snn = (sqf[i] / (baver * baver)) / ((double)natoms * nqi);"),
    "snn_without_baver": ("double snn = (sqf[i] / (baver * baver))", "double snn = (sqf[i])"),
    "s11_N_for_N1": ("double s11 = x1 * sqfp[0 * ns + i] / ((double)ntype[0]",
                     "double s11 = x1 * sqfp[0 * ns + i] / ((double)natoms"),
    "scc_swapped_fractions": ("double scc = x2 * x2 * s11 + x1 * x1 * s22", "double scc = x1 * x1 * s11 + x2 * x2 * s22"),
    "gxy_without_zgrid_Lz": ("f * v * (zgrid / box[2]) * hxy", "f * v * hxy"),
    "gxy_without_pair_factor": ("double f = (j == l) ? 2.0 : 1.0;", "double f = 1.0;"),
    "gxy_without_nconf": ("(dv * (double)nconf);", "(dv);"),
    "member_wrong_q2": ("q2 = kf1 * kf1 + kf2 * kf2 + kf3 * kf3;", "q2 = kf1 * kf1 + kf2 * kf2 + kf2 * kf3;"),
    "member_wrong_rounding": ("int ind = (int)(q / g->dq + 0.501f);", "int ind = (int)(q / g->dq + 0.4f);"),
}


@pytest.mark.parametrize("name", sorted(MUTATIONS))
def test_pins_fail_on_planted_mistake(name, tmp_path):
    old, new = MUTATIONS[name]
    src = open(SRC).read()
    assert src.count(old) == 1, f"{name}: the line it plants into is not in the oracle"
    mut = tmp_path / "trj_oracle.c"
    mut.write_text(src.replace(old, new))
    (tmp_path / "trj_oracle.h").write_text(open(os.path.join(ROOT, "oracle", "trj_oracle.h")).read())
    lib = tmp_path / "libtrj_oracle_mut.so"
    subprocess.run(["gcc", "-O2", "-fPIC", "-std=gnu11", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-shared", "-o", str(lib),
                    str(mut), "-lm"], check=True)
    env = dict(os.environ, TRJ_ORACLE_LIB=str(lib))
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                        os.path.join(ROOT, "tests", "test_oracle_pins.py")], cwd=ROOT, env=env,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 1, f"{name}: the pins passed with the planted mistake\n{r.stdout[-2000:]}"
