"""Pins of the oracle's charge-weighted functions (g_qq and S_qq) against
closed forms and the unweighted functions they must reduce to."""
import math

import numpy as np
import pytest

from oracle import oracle as O


def _frame(n=600, L=11.0, seed=3, nsp=2):
    rng = np.random.default_rng(seed)
    pos = rng.random((n, 3)) * L
    it = rng.integers(1, nsp + 1, n).astype(np.int32)
    return pos, it, [L] * 3


def test_qq_scale_powers_of_two():
    assert O.qq_scale([0.0, 0.0]) == 2.0 ** 30
    assert O.qq_scale([1.0, -1.0]) == 2.0 ** 29          # 1 < 2^1
    assert O.qq_scale([0.5]) == 2.0 ** 31                # 0.25 < 2^-1
    assert O.qq_scale([0.3, -0.1]) == 2.0 ** 33          # 0.09 < 2^-3
    assert O.qq_scale([2.0]) == 2.0 ** 27                # 4 < 2^3


def test_unit_charges_are_pair_counts():
    pos, it, box = _frame()
    rc, dr = 4.0, 0.05
    counts = O.rdf_counts(pos, it, 3, 2, box, (1, 1, 1), rc, dr, method="brute")
    fixed, scale = O.rdf_qq_fixed(pos, np.ones(len(it)), 3, box, (1, 1, 1), rc, dr)
    assert scale == 2.0 ** 29
    np.testing.assert_array_equal(fixed, counts.sum(axis=0) * 2 ** 29)


def test_species_sign_charges():
    pos, it, box = _frame(seed=4)
    rc, dr = 4.0, 0.05
    counts = O.rdf_counts(pos, it, 3, 2, box, (1, 1, 1), rc, dr, method="brute")  # rows 11, 12, 22
    q = np.where(it == 1, 1.0, -1.0)
    fixed, scale = O.rdf_qq_fixed(pos, q, 3, box, (1, 1, 1), rc, dr)
    np.testing.assert_array_equal(fixed, (counts[0] + counts[2] - counts[1]) * 2 ** 29)
    # q -> 2q: scale drops by 4, the fixed-point sums are unchanged
    fixed2, scale2 = O.rdf_qq_fixed(pos, 2 * q, 3, box, (1, 1, 1), rc, dr)
    assert scale2 == scale / 4
    np.testing.assert_array_equal(fixed2, fixed)


def test_two_atoms_closed_form():
    L = 10.0
    pos = np.array([[1.0, 1.0, 1.0], [1.0, 1.0, 3.37]])
    q = np.array([0.413, -0.731])
    dr = 0.1
    fixed, scale = O.rdf_qq_fixed(pos, q, 3, [L] * 3, (1, 1, 1), 4.0, dr)
    k = int(math.floor(2.37 / dr + 0.5))
    want = np.zeros_like(fixed)
    want[k - 1] = np.rint(q[0] * q[1] * scale)
    np.testing.assert_array_equal(fixed, want)


def test_accumulate_and_gqq():
    fixed = np.array([0, 2 ** 29, -3 * 2 ** 29, 2 ** 30], dtype=np.int64)
    q = np.array([1.0, -1.0, 1.0, -1.0])  # avq2 = 1
    acc = O.qq_accumulate(fixed, 2.0 ** 29, q, np.zeros(4))
    np.testing.assert_array_equal(acc, [0.0, 1.0, -3.0, 2.0])
    q2 = np.array([2.0, 0.0])  # avq2 = 2
    acc = O.qq_accumulate(fixed, 2.0 ** 29, q2, acc)
    np.testing.assert_array_equal(acc, [0.0, 1.5, -4.5, 3.0])
    # g_qq has gmix's form with N_j N_l -> N^2 and the factor 2 of a like pair
    lsmax, dr, V, N, nconf = 6, 0.1, 1000.0, 50, 3
    a = np.array([5.0, 7.0, 11.0, 2.0, 0.0, 0.0])
    g = O.rdf_gqq(3, lsmax, dr, V, N, nconf, a)
    counts = np.zeros((1, lsmax), dtype=np.int64)
    counts[0] = a.astype(np.int64)
    gm = O.rdf_gmix(3, 1, lsmax, dr, V, [N], nconf, counts)
    np.testing.assert_allclose(g, gm[:, 1], rtol=1e-15)


def test_sq_frame_q_reductions():
    pos, it, box = _frame(n=300, seed=5)
    g = O.sq_setup(3, box, 6.0, 6.0)
    w = O.wrap(pos, box, (1, 1, 1))
    sqf, _ = O.sq_frame(w, it, 3, 2, box, g, [1.0, 1.0])
    sqfq = O.sq_frame_q(w, np.ones(len(it)), 3, box, g)
    np.testing.assert_allclose(sqfq, sqf, rtol=1e-12, atol=1e-9)
    sqfq3 = O.sq_frame_q(w, np.full(len(it), -0.3), 3, box, g)
    np.testing.assert_allclose(sqfq3, 0.09 * sqf, rtol=1e-12, atol=1e-9)
    one = O.sq_frame_q(w[:1], np.array([0.7]), 3, box, g)
    nq = O.sq_nq(3, g)
    # one atom: |q e^{i phi}|^2 = q^2 per summed vector
    summed = np.zeros(g.nqmax + 1)
    summed = O.sq_frame_q(w[:1], np.array([1.0]), 3, box, g, summed)
    np.testing.assert_allclose(one, 0.49 * summed, rtol=1e-14)
    assert np.all(np.abs(summed - np.rint(summed)) < 1e-9) and summed.sum() > 0
    assert nq.sum() > 0
