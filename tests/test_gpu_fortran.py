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


REF_EXE = os.path.join(ROOT, "trj_analysis_amd", "fortran", "reference_api", "ref_driver")


def test_reference_signature_modules_match_oracle(tmp_path):
    """The reference's own procedures (RDF_init(nsp), RDFcomp(Nsites, Iconf,
    nbcuda, nthread, units), sq_init(Nsites, nsp, nbcuda), SQcalc(),
    sq_transfer_gpu_cpu()) on the module variables of mod_common / mod_input,
    driven by the reference's frame loop (trj_analysis.f90:193-217): histomix
    exact (upper triangle, as rdf.cuf:301-310 fills it), gtot from the
    reference's formula (rdf.cuf:311-341), rmingr, the S(Q) plan (nqmax,
    nqmin, dq, the moved qmin) and shells against the oracle."""
    frames = list(W.lj_binary_4k(nframes=3))
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
            fr.pos.astype(np.float32).tofile(f)
    print("case: reference-signature Fortran driver", flush=True)
    subprocess.run([REF_EXE, str(fin), str(fout)], check=True, timeout=120)
    raw = open(fout, "rb").read()
    lsmax, nit, nqmax, nqmin = np.frombuffer(raw[:16], dtype=np.int32)
    off = 16

    def take(count, dtype):
        nonlocal off
        a = np.frombuffer(raw[off:off + np.dtype(dtype).itemsize * count], dtype=dtype)
        off += np.dtype(dtype).itemsize * count
        return a

    h = take(lsmax * nsp * nsp, np.float32).reshape(nsp, nsp, lsmax)  # h[l, k, :] = histomix(:, k, l)
    gtot = take(lsmax + 1, np.float32)
    rmingr, qmin_eff, dq = take(3, np.float32)
    sqf = take(nqmax + 1, np.float32)
    sqfp = take((nqmax + 1) * nsp, np.float32).reshape(nsp, nqmax + 1)
    nq = take(nqmax + 1, np.int32)
    assert off == len(raw)

    box32 = frames[0].box.astype(np.float32).astype(np.float64)
    ls = O.rdf_lsmax(3, box32, float(rcrdf), float(deltar))
    want = sum(O.rdf_counts(fr.pos.astype(np.float32), fr.itype, 3, 2, box32, (1, 1, 1), float(rcrdf),
                            float(deltar), lsmax=ls, method="cells") for fr in frames)
    assert lsmax == ls and nit == 3
    np.testing.assert_array_equal(np.stack([h[0, 0], h[1, 0], h[1, 1]]), want.astype(np.float32))
    np.testing.assert_array_equal(h[0, 1], np.zeros(lsmax, np.float32))  # histomix(:, 2, 1) stays 0
    # gtot(i) = 2 V sum_{k<=l} histomix(i,k,l) / (deltaV_i natms^2), i = 1..lsmax
    vol = float(np.prod(box32))
    ri = np.arange(1, lsmax + 1) * float(deltar)
    dv = 4 * np.pi * ((ri + float(deltar) / 2) ** 3 - (ri - float(deltar) / 2) ** 3) / 3
    np.testing.assert_allclose(gtot[1:], 2 * vol * want.sum(axis=0) / (dv * n * n), rtol=2e-5)
    assert abs(rmingr - O.first_minimum(3, 2, ls, float(deltar), vol, n, want)) < 1e-4
    g = O.sq_setup(3, box32, float(qmax), float(qmin))
    assert (nqmax, nqmin) == (g.nqmax, g.nqmin)
    assert abs(dq - g.dq) <= 1e-6 * g.dq and abs(qmin_eff - g.qminp) <= 1e-5 * g.qminp
    np.testing.assert_array_equal(nq, O.sq_nq(3, g).astype(np.int32))
    osqf = np.zeros(g.nqmax + 1)
    osqfp = np.zeros((2, g.nqmax + 1))
    for fr in frames:
        O.sq_frame(O.wrap(fr.pos.astype(np.float32), box32, (1, 1, 1)), fr.itype, 3, 2, box32, g,
                   bsc.astype(np.float64), osqf, osqfp)
    m = osqf > 0
    assert np.max(np.abs(sqf[m] - osqf[m]) / osqf[m]) < 1e-6  # real(4) of the FP64 sums
    m = osqfp > 0
    assert np.max(np.abs(sqfp[m] - osqfp[m]) / osqfp[m]) < 1e-6
