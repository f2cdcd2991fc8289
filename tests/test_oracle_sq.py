"""Pins of the S(Q) half of the oracle: closed forms, invariants, and the
agreement of its q-set with the library's host plan."""
import math

import numpy as np
import pytest

from oracle import oracle as O


def _kv(rng, n, lo=-6, hi=7):
    return rng.integers(lo, hi, size=(n, 3)).astype(np.int32)


def test_single_atom_unit_modulus():
    rng = np.random.default_rng(1)
    box = [7.3, 8.1, 9.9]
    pos = np.array([[1.234, 5.678, 9.1]])
    rho = O.sq_rho(pos, [1], 1, box, _kv(rng, 50))
    np.testing.assert_allclose(np.abs(rho[:, 0]) ** 2, 1.0, rtol=0, atol=1e-15)
    # phase itself: exp(i 2 pi n.r/L)
    kv = np.array([[1, 2, 3]])
    want = np.exp(2j * np.pi * (1 * 1.234 / 7.3 + 2 * 5.678 / 8.1 + 3 * 9.1 / 9.9))
    assert abs(O.sq_rho(pos, [1], 1, box, kv)[0, 0] - want) < 1e-14


def test_two_atoms_cosine():
    rng = np.random.default_rng(2)
    box = np.array([10.0, 11.0, 12.0])
    pos = np.array([[0.5, 1.5, 2.5], [3.25, 7.75, 11.5]])
    kv = _kv(rng, 40)
    rho = O.sq_rho(pos, [1, 1], 1, box, kv)[:, 0]
    q = 2 * np.pi * kv / box
    want = 2 + 2 * np.cos(q @ (pos[1] - pos[0]))
    np.testing.assert_allclose(np.abs(rho) ** 2, want, atol=1e-13)
    # species split: two species, one atom each -> |rho_s|^2 = 1
    r2 = O.sq_rho(pos, [1, 2], 2, box, kv)
    np.testing.assert_allclose(np.abs(r2) ** 2, 1.0, atol=1e-14)


def test_perfect_lattice_bragg():
    m, a = 4, 1.7
    L = m * a
    g = np.arange(m) * a + 0.3
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    n = len(pos)
    kv = np.array([[4, 0, 0], [0, 8, 4], [1, 0, 0], [2, 3, 1], [4, 4, 1]], dtype=np.int32)
    rho = O.sq_rho(pos, np.ones(n, np.int32), 1, [L] * 3, kv)[:, 0]
    assert abs(abs(rho[0]) - n) < 1e-10 and abs(abs(rho[1]) - n) < 1e-10
    np.testing.assert_allclose(np.abs(rho[2:]), 0.0, atol=1e-10)


def test_translation_invariance_and_inversion():
    rng = np.random.default_rng(3)
    box = np.array([9.0, 9.0, 9.0])
    pos = rng.random((300, 3)) * box
    it = rng.integers(1, 3, 300).astype(np.int32)
    kv = _kv(rng, 30)
    a = np.abs(O.sq_rho(pos, it, 2, box, kv)) ** 2
    b = np.abs(O.sq_rho((pos + [1.1, 2.2, 3.3]) % box, it, 2, box, kv)) ** 2
    c = np.abs(O.sq_rho(pos, it, 2, box, -kv)) ** 2
    np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(a, c, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("ndim,L,qmax,qmin", [(3, 12.0, 6.0, 3.0), (3, 17.0998, 10.0, 10.0),
                                              (3, 9.0, 4.0, 0.4), (2, 40.0, 3.0, 1.5)])
def test_single_atom_frame_counts_summed_vectors(ndim, L, qmax, qmin):
    """With one atom, sqfp[shell] is the number of vectors summed in the shell:
    ties or_sq_member and the axis set to the shell bookkeeping."""
    box = [L, L, L] if ndim == 3 else [L, L, 1.0]
    g = O.sq_setup(ndim, box, qmax, qmin)
    pos = np.array([[0.3, 0.7, 0.1 if ndim == 3 else 0.0]])
    sqf, sqfp = O.sq_frame(pos, [1], ndim, 1, box, g, [1.0])
    want = np.zeros(g.nqmax + 1)
    if g.nqmin > 1:
        kz = g.kext[2] if ndim == 3 else 1
        for a in range(g.kext[0]):
            for b in range(g.kext[1]):
                for c in range(kz):
                    s = O.sq_member(ndim, g, a, b, c)
                    if s >= 0:
                        want[s] += 1
        k0 = g.nqmin + 1
    else:
        k0 = 1
    for k in range(k0, g.nqmax + 1):
        want[k] += 3 if ndim == 3 else 2
    np.testing.assert_allclose(sqfp[0], want, atol=1e-9)
    np.testing.assert_allclose(sqf, want, atol=1e-9)
    # the library's consistent count is the same set
    from trj_analysis_amd import trjhip as T
    p = T.sq_params(ndim, 1, qmax, qmin, flags=T.TRJ_SQ_NQ_CONSISTENT)
    info, nq = T.sq_plan(p, box)
    np.testing.assert_array_equal(nq, want.astype(np.int64))
    p = T.sq_params(ndim, 1, qmax, qmin)
    info, nq = T.sq_plan(p, box)
    np.testing.assert_array_equal(nq, O.sq_nq(ndim, g))


def test_member_shells_brute():
    """or_sq_member against a brute enumeration of the FP32 rule on a tiny grid."""
    g = O.sq_setup(3, [11.0, 11.0, 11.0], 8.0, 5.0)
    f32 = np.float32
    for a in range(g.kext[0]):
        for b in range(g.kext[1]):
            for c in range(g.kext[2]):
                kf = [f32(a) * f32(g.fk[0]), f32(b) * f32(g.fk[1]), f32(c) * f32(g.fk[2])]
                q2 = f32(f32(kf[0] * kf[0]) + f32(kf[1] * kf[1])) + f32(kf[2] * kf[2])
                if q2 > 0 and q2 <= f32(g.qmin2):
                    q = np.sqrt(q2, dtype=np.float32)
                    s = int(f32(f32(q / f32(g.dq)) + f32(0.501)))
                    want = s if s <= g.nqmax else -1
                else:
                    want = -1
                assert O.sq_member(3, g, a, b, c) == want


def test_columns_closed_form():
    nsp, nqmax = 2, 4
    g = O.sq_setup(3, [10.0] * 3, 2.0, 2.0)
    g.nqmax = nqmax
    ntype = [30, 70]
    n = 100
    sqf = np.array([0, 10.0, 20.0, 30.0, 40.0])
    sqfp = np.array([[0, 3.0, 4.0, 5.0, 6.0], [0, 7.0, 8.0, 9.0, 10.0]])
    nq = np.array([0, 6, 12, 8, 6])
    cols, mix = O.sq_columns(nsp, g, n, ntype, [1.0, 1.0], 5, sqf, sqfp, nq)
    assert cols.shape == (nqmax, 7)
    for i in range(1, nqmax + 1):
        snn = sqf[i] / (n * 5 * nq[i])
        s11 = 0.3 * sqfp[0, i] / (30 * 5 * nq[i])
        s22 = 0.7 * sqfp[1, i] / (70 * 5 * nq[i])
        s12 = 0.5 * (snn - s11 - s22)
        assert cols[i - 1, 1] == pytest.approx(snn)
        assert cols[i - 1, 1] == pytest.approx(s11 + s22 + 2 * s12)
        assert cols[i - 1, 2] == pytest.approx(0.49 * s11 + 0.09 * s22 - 2 * 0.21 * s12)
        assert mix[i - 1, 1] == pytest.approx(sqfp[0, i] / (30 * 5 * nq[i]))
    # weighted: sqf is divided by <b>^2
    cols, _ = O.sq_columns(nsp, g, n, ntype, [2.0, 0.5], 5, sqf, sqfp, nq)
    assert cols.shape == (nqmax, 3)
    bav = (30 * 2.0 + 70 * 0.5) / 100
    assert cols[0, 1] == pytest.approx(sqf[1] / bav ** 2 / (n * 5 * nq[1]))


def test_f64_timing_port_equals_long_double_oracle():
    """or_sq_frame_f64 (FP64 sincos, the CPU baseline's timing port) sums the
    same set into the same shells as or_sq_frame: equal to 1e-12 relative."""
    rng = np.random.default_rng(9)
    L = 12.0
    pos = rng.random((900, 3)) * L
    it = rng.integers(1, 3, 900).astype(np.int32)
    g = O.sq_setup(3, [L] * 3, 7.0, 7.0)
    a, ap = O.sq_frame(pos, it, 3, 2, [L] * 3, g, [1.0, 0.5])
    b, bp = O.sq_frame_f64(pos, it, 3, 2, [L] * 3, g, [1.0, 0.5])
    m = a > 0
    assert np.max(np.abs(b[m] - a[m]) / a[m]) < 1e-12
    mp = ap > 0
    assert np.max(np.abs(bp[mp] - ap[mp]) / ap[mp]) < 1e-12


def test_f64_rho_port_equals_long_double_oracle():
    rng = np.random.default_rng(10)
    box = np.array([9.0, 10.0, 11.0])
    pos = rng.random((500, 3)) * box
    it = rng.integers(1, 4, 500).astype(np.int32)
    kv = _kv(rng, 60, -9, 10)
    a = O.sq_rho(pos, it, 3, box, kv)
    b = O.sq_rho_f64(pos, it, 3, box, kv)
    assert np.max(np.abs(a - b)) < 1e-11
