"""Pins of the oracle's confined-slice functions (rdfxy, sqf3xy/sqf2xy):
hand cases of the FP32 slice rule, independent numpy pair counts, and the
reductions to the unsliced oracle."""
import numpy as np
import pytest

from oracle import oracle as O


def test_slice_rule_hand_cases():
    zs = [0.0, 4.0]
    assert O.slice_of(1.0, zs, 8.0) == 0
    assert O.slice_of(2.0, zs, 8.0) == 0          # equidistant: the first (minloc)
    assert O.slice_of(3.5, zs, 8.0) == 1
    assert O.slice_of(-4.0, zs, 8.0) == 0         # |d| = zgrid/2 is inside
    assert O.slice_of(-4.5, zs, 8.0) == -1
    # FP32: 1.5 + 1e-9 rounds to 1.5, inside zgrid/2 = 0.5 of 1.0
    assert O.slice_of(1.5 + 1e-9, [1.0], 1.0) == 0
    assert O.slice_of(1.5 + 1e-6, [1.0], 1.0) == -1


def _frame(n=700, L=14.0, H=20.0, seed=1, nsp=2):
    rng = np.random.default_rng(seed)
    pos = np.empty((n, 3))
    pos[:, :2] = rng.random((n, 2)) * L
    pos[:, 2] = rng.random(n) * H - 5.0
    it = rng.integers(1, nsp + 1, n).astype(np.int32)
    return pos, it, [L, L, H + 8.0]


def _xy_d2(pos, L):
    d = pos[None, :, :2] - pos[:, None, :2]
    d -= L * np.round(d / L)
    return (d ** 2).sum(-1)


def test_rdf_xy_one_slice_counts_all_pairs_by_floor_bins():
    pos, it, box = _frame()
    rc, dr = 5.0, 0.1
    counts, count, countp, _, _ = O.rdf_xy_counts(pos, it, 2, box, rc, dr, [5.0], 1e6)
    d2 = _xy_d2(pos, box[0])
    iu = np.triu_indices(len(it), 1)
    r = np.sqrt(d2[iu])
    r = r[r < rc]
    ls = counts.shape[2]
    want, _ = np.histogram(r, bins=np.arange(ls + 1) * dr)
    np.testing.assert_array_equal(counts.sum(axis=(0, 1)), want)
    # species split: pairs (1,1), (1,2), (2,2) in ind_ij order
    ti, tj = it[iu[0]], it[iu[1]]
    sel = np.sqrt(d2[iu]) < rc
    assert counts[0, 0].sum() == np.sum(sel & (ti == 1) & (tj == 1))
    assert counts[0, 1].sum() == np.sum(sel & (ti != tj))
    assert counts[0, 2].sum() == np.sum(sel & (ti == 2) & (tj == 2))
    # counts leave out the last atom of the species-sorted order
    last = np.max(np.nonzero(it == 2)[0])
    assert count[0] == len(it) - 1
    assert countp[0, 1] == np.sum(it == 2) - 1 and countp[0, 0] == np.sum(it == 1)
    assert last >= 0


def test_rdf_xy_slices_partition_pairs():
    pos, it, box = _frame(seed=2)
    rc, dr = 4.0, 0.05
    zs, zg = [0.0, 8.0], 6.0
    counts, count, countp, _, _ = O.rdf_xy_counts(pos, it, 2, box, rc, dr, zs, zg)
    sl = np.array([O.slice_of(z, zs, zg) for z in pos[:, 2]])
    d2 = _xy_d2(pos, box[0])
    for s in range(2):
        m = np.nonzero(sl == s)[0]
        sub = d2[np.ix_(m, m)][np.triu_indices(len(m), 1)]
        assert counts[s].sum() == np.sum(sub < rc * rc)
    assert count.sum() + 1 >= np.sum(sl >= 0) >= count.sum()


def test_rdf_xy_two_atoms_and_charges():
    pos = np.array([[1.0, 1.0, 0.2], [1.0, 13.7, -0.3], [5.0, 5.0, 9.0]])  # image in y: r = 1.3
    it = np.array([1, 2, 1], dtype=np.int32)
    q = np.array([0.7, -0.4, 1.1])
    counts, count, countp, fixed, scale = O.rdf_xy_counts(pos, it, 2, [14.0, 14.0, 20.0], 4.0, 0.25,
                                                         [0.0], 2.0, q=q)
    k = int(1.3 / 0.25) + 1                      # floor rule: bin 6
    assert counts[0, 1, k - 1] == 1 and counts.sum() == 1
    assert fixed[0, k - 1] == np.rint(0.7 * -0.4 * scale) and fixed.sum() == fixed[0, k - 1]
    # atom 2 (species 2) is the last of the species-sorted order: not counted
    assert list(count) == [1] and list(countp[0]) == [1, 0]


def test_rdf_xy_accumulate_hand():
    counts = np.zeros((1, 3, 4), dtype=np.int64)
    counts[0, 0, 1] = 6
    counts[0, 1, 2] = 10
    counts[0, 2, 3] = 3
    h = O.rdf_xy_accumulate(2, counts, np.array([[3, 5]]))
    assert h[0, 0, 0, 1] == 6 / 9 and h[0, 0, 1, 2] == 10 / 15 and h[0, 1, 1, 3] == 3 / 25
    h0 = O.rdf_xy_accumulate(2, counts, np.array([[0, 5]]))
    assert h0[0, 0, 0].sum() == 0 and h0[0, 0, 1].sum() == 0 and h0[0, 1, 1, 3] == 3 / 25


def test_sq_xy_one_slice_is_2d_sq_over_n():
    pos, it, box = _frame(n=300, L=12.0, seed=3)
    g = O.sq_setup(2, box, 12.0, 10.5)            # l = 20 -> 8 after the 16-adjust
    assert g.nqmin > 1                            # both sets start the axes at nqmin+1
    sqf, sqfp, sqfq = O.sq_frame_xy(pos, it, 2, box, g, [1.0, 0.6], [5.0], 1e6)
    pos2 = pos.copy()
    pos2[:, 2] = 0.0
    osqf, osqfp = O.sq_frame(pos2, it, 2, 2, box, g, [1.0, 0.6])
    np.testing.assert_allclose(sqf[0], osqf / 300, rtol=1e-13)
    ntype = np.bincount(it)[1:]
    np.testing.assert_allclose(sqfp[0], osqfp / ntype[:, None], rtol=1e-13)
    assert not sqfq.any()


def test_sq_xy_single_atom_slices_and_unit_charges():
    pos = np.array([[1.0, 2.0, 0.0], [3.0, 7.0, 10.0]])
    it = np.array([1, 2], dtype=np.int32)
    box = [11.0, 11.0, 20.0]
    g = O.sq_setup(2, box, 5.0, 3.0)
    sqf, sqfp, sqfq = O.sq_frame_xy(pos, it, 2, box, g, [1.0, 1.0], [0.0, 10.0], 2.0, q=np.ones(2))
    # one atom per slice: |rho|^2 / 1 = 1 per vector, so shells count the summed vectors
    per_shell = sqf[0]
    assert np.allclose(sqf[1], per_shell) and np.allclose(sqfq, sqf)
    assert np.allclose(sqfp[0, 0], per_shell) and not sqfp[0, 1].any() and np.allclose(sqfp[1, 1], per_shell)
    assert np.all(np.abs(per_shell - np.rint(per_shell)) < 1e-12) and per_shell.sum() > 0
    assert g.nqmin < 1 and per_shell[1] == 2      # axis vectors (1,0), (0,1) from k = 1


def test_sq_xy_axis_start_when_nqmin_is_one():
    """qmin below fk/2: l = 0, nqmin = 1 (sq.cuf:196-201); sqf2xy then starts
    at k = nqmin + 1 = 2, so shell 1 gets no vector (nq(1) = 2 regardless)."""
    pos = np.array([[1.0, 2.0, 0.0]])
    it = np.array([1], dtype=np.int32)
    box = [11.0, 11.0, 20.0]
    g = O.sq_setup(2, box, 5.0, 0.2)
    assert g.nqmin == 1
    sqf, _, _ = O.sq_frame_xy(pos, it, 1, box, g, [1.0], [0.0], 2.0)
    assert sqf[0, 1] == 0.0 and np.allclose(sqf[0, 2:], 2.0)
    assert list(O.sq_nq(2, g)[1:3]) == [2, 2]
