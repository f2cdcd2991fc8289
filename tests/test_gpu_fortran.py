"""The Fortran host path on the GPU: the reference's frame loop (RDF_init,
RDFcomp, sq_init, SQcalc, sq_transfer_gpu_cpu) through the ISO_C_BINDING
modules, checked against the oracle."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from trj_analysis_amd import workloads as W

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "trj_analysis_amd", "fortran", "example_driver")


def test_fortran_driver_matches_oracle(tmp_path):
    frames = list(W.lj_binary_4k(nframes=2))
    nsp, ndim, n = 2, 3, len(frames[0].itype)
    deltar, rcrdf, qmax, qmin = np.float32(0.02), np.float32(4.0), np.float32(10.0), np.float32(10.0)
    bsc = np.array([1.0, 0.5], dtype=np.float32)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        np.array([ndim, nsp, n, len(frames)], dtype=np.int32).tofile(f)
        np.array([deltar, rcrdf, qmax, qmin], dtype=np.float32).tofile(f)
        np.array([1, 1, 1], dtype=np.int32).tofile(f)
        bsc.tofile(f)
        for fr in frames:
            fr.box.astype(np.float32).tofile(f)
            fr.itype.astype(np.int32).tofile(f)
            fr.pos.astype(np.float32).tofile(f)  # (N,3) C order == r(3,N) Fortran order
    print("case: fortran driver", flush=True)
    subprocess.run([EXE, str(fin), str(fout)], check=True, timeout=120)
    raw = open(fout, "rb").read()
    lsmax, nit, nqmax = np.frombuffer(raw[:12], dtype=np.int32)
    off = 12
    h = np.frombuffer(raw[off:off + 8 * lsmax * nsp * nsp], dtype=np.int64).reshape(nsp, nsp, lsmax)
    off += 8 * lsmax * nsp * nsp
    sqf = np.frombuffer(raw[off:off + 8 * (nqmax + 1)], dtype=np.float64)
    off += 8 * (nqmax + 1)
    sqfp = np.frombuffer(raw[off:off + 8 * (nqmax + 1) * nsp], dtype=np.float64).reshape(nsp, nqmax + 1)
    off += 8 * (nqmax + 1) * nsp
    nq = np.frombuffer(raw[off:off + 8 * (nqmax + 1)], dtype=np.int64)
    # oracle on the same FP32 parameters
    box32 = frames[0].box.astype(np.float32).astype(np.float64)
    ls = O.rdf_lsmax(3, box32, float(rcrdf), float(deltar))
    want = sum(O.rdf_counts(fr.pos.astype(np.float32), fr.itype, 3, 2, box32, (1, 1, 1), float(rcrdf),
                            float(deltar), lsmax=ls, method="brute") for fr in frames)
    assert lsmax == ls
    got = np.stack([h[0, 0], h[0, 1], h[1, 1]])
    np.testing.assert_array_equal(got, want)
    g = O.sq_setup(3, box32, float(qmax), float(qmin))
    osqf = np.zeros(g.nqmax + 1)
    osqfp = np.zeros((2, g.nqmax + 1))
    for fr in frames:
        O.sq_frame(O.wrap(fr.pos.astype(np.float32), box32, (1, 1, 1)), fr.itype, 3, 2, box32, g,
                   bsc.astype(np.float64), osqf, osqfp)
    m = osqf > 0
    assert np.max(np.abs(sqf[m] - osqf[m]) / osqf[m]) < 1e-12
    m = osqfp > 0
    assert np.max(np.abs(sqfp[m] - osqfp[m]) / osqfp[m]) < 1e-12
    np.testing.assert_array_equal(nq, O.sq_nq(3, g))
