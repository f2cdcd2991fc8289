"""Pins of the RDF half of the oracle: closed forms and invariants that a
dropped term, a wrong index or a wrong image would fail."""
import math

import numpy as np
import pytest

from oracle import oracle as O


def test_ind_ij_closed_form():
    # upper-triangle row-major enumeration (1,1),(1,2),..,(1,n),(2,2),..
    for nsp in range(1, 11):
        k = 0
        for i in range(1, nsp + 1):
            for j in range(i, nsp + 1):
                k += 1
                assert O.ind_ij(i, j, nsp) == k
                assert O.ind_ij(j, i, nsp) == k
        assert k == nsp * (nsp + 1) // 2


def test_nint_rounds_half_away():
    assert O.nint(2.5) == 3 and O.nint(-2.5) == -3 and O.nint(2.4999) == 2


def test_geometry_cap_and_nonperiodic():
    g = O.geometry(3, [10.0, 12.0, 8.0], [1, 1, 0], 5.0)
    assert g.rcrdf == 4.0 and g.side2 == 16.0
    assert list(g.lmi) == [10.0, 12.0, 20.0]
    assert g.volume == 960.0
    g = O.geometry(3, [10.0, 12.0, 8.0], [1, 1, 1], -1.0)
    assert g.side2 == 16.0  # rcrdf <= 0 -> (side/2)^2


@pytest.mark.parametrize("x", [-1e-17, -0.0, 0.0, 9.999999999999999, 10.0, -10.0, 23.7, -23.7])
def test_wrap_invariants(x):
    w = O.lib().or_wrap1(x, 10.0)
    assert 0.0 <= w < 10.0
    k = round((x - w) / 10.0)
    assert abs((x - w) - 10.0 * k) < 1e-12


def test_two_atoms_closed_form():
    L, dr = 20.0, 0.05
    box = [L, L, L]
    for d in [0.73, 1.0, 3.3, 4.9]:
        for shift in [0.0, 19.5]:  # second case crosses the periodic boundary
            pos = np.array([[shift, 1.0, 2.0], [shift + d, 1.0, 2.0]])
            pos[:, 0] %= L
            c = O.rdf_counts(pos, [1, 2], 3, 2, box, (1, 1, 1), 5.0, dr, method="brute")
            k = O.nint(d / dr)
            assert c.sum() == 1 and c[1, k - 1] == 1  # pair (1,2) is ij=2


def test_simple_cubic_shells():
    """m^3 lattice, spacing 1: pairs at |n|^2 = s are N*r3(s)/2 (r3 = OEIS A005875)."""
    r3 = {1: 6, 2: 12, 3: 8, 4: 6, 5: 24, 6: 24}
    m, dr, rc = 8, 0.05, 2.6
    g = np.arange(m, dtype=np.float64) + 0.37  # a translation off the grid
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    n = len(pos)
    itype = np.ones(n, dtype=np.int32)
    for method in ["brute", "cells"]:
        c = O.rdf_counts(pos, itype, 3, 1, [m, m, m], (1, 1, 1), rc, dr, method=method)[0]
        want = np.zeros_like(c)
        for s, r in r3.items():
            want[O.nint(math.sqrt(s) / dr) - 1] += n * r // 2
        np.testing.assert_array_equal(c, want)


def _random_frame(rng, n, L, nsp, dtype=np.float64):
    pos = (rng.random((n, 3)) * L).astype(dtype)
    itype = rng.integers(1, nsp + 1, size=n).astype(np.int32)
    return pos, itype


@pytest.mark.parametrize("L,rc,dr,per", [(9.0, 4.0, 0.02, (1, 1, 1)), (15.0, 3.1, 0.05, (1, 1, 1)),
                                         (12.0, 2.5, 0.03, (1, 1, 0)), (7.0, 3.5, 0.1, (0, 1, 1))])
def test_cells_equal_brute(L, rc, dr, per):
    rng = np.random.default_rng(7)
    pos, itype = _random_frame(rng, 1500, L, 3)
    a = O.rdf_counts(pos, itype, 3, 3, [L, L, L], per, rc, dr, method="brute")
    b = O.rdf_counts(pos, itype, 3, 3, [L, L, L], per, rc, dr, method="cells")
    np.testing.assert_array_equal(a, b)


def test_2d_cells_equal_brute():
    rng = np.random.default_rng(3)
    pos, itype = _random_frame(rng, 1200, 30.0, 2)
    pos[:, 2] = 0.0
    a = O.rdf_counts(pos, itype, 2, 2, [30.0, 30.0, 1.0], (1, 1, 1), 6.0, 0.05, method="brute")
    b = O.rdf_counts(pos, itype, 2, 2, [30.0, 30.0, 1.0], (1, 1, 1), 6.0, 0.05, method="cells")
    np.testing.assert_array_equal(a, b)


def test_permutation_translation_invariance():
    rng = np.random.default_rng(11)
    L = 11.0
    pos, itype = _random_frame(rng, 800, L, 2)
    base = O.rdf_counts(pos, itype, 3, 2, [L] * 3, (1, 1, 1), 4.0, 0.02, method="brute")
    perm = rng.permutation(len(pos))
    np.testing.assert_array_equal(
        base, O.rdf_counts(pos[perm], itype[perm], 3, 2, [L] * 3, (1, 1, 1), 4.0, 0.02, method="brute"))
    # translation by a multiple of the bin-safe amount: integer lattice shift
    np.testing.assert_array_equal(
        base, O.rdf_counts(pos + L, itype, 3, 2, [L] * 3, (1, 1, 1), 4.0, 0.02, method="brute"))


def test_total_pairs_in_range():
    """Sum of counts = pairs with dr/2 <= r < rc (bins 1..lsmax), independent count."""
    rng = np.random.default_rng(5)
    L, rc, dr = 10.0, 4.0, 0.02
    pos, itype = _random_frame(rng, 700, L, 1)
    c = O.rdf_counts(pos, itype, 3, 1, [L] * 3, (1, 1, 1), rc, dr, method="cells")
    d = pos[:, None, :] - pos[None, :, :]
    d -= L * np.round(d / L)
    r = np.sqrt((d ** 2).sum(-1))[np.triu_indices(len(pos), 1)]
    assert c.sum() == np.count_nonzero((r < rc) & (r >= dr / 2))


def test_gmix_closed_form():
    nsp, lsmax, dr, vol = 2, 6, 0.5, 1000.0
    counts = np.arange(1, 3 * lsmax + 1, dtype=np.int64).reshape(3, lsmax)
    out = O.rdf_gmix(3, nsp, lsmax, dr, vol, [10, 20], 4, counts)
    assert out.shape == (lsmax - 2, 4)
    for i in range(1, lsmax - 1):
        r = i * dr
        dv = 4 * math.pi / 3 * ((r + dr / 2) ** 3 - (r - dr / 2) ** 3)
        assert out[i - 1, 0] == pytest.approx(r)
        assert out[i - 1, 1] == pytest.approx(2 * vol * counts[0, i - 1] / (dv * 10 * 10 * 4))
        assert out[i - 1, 2] == pytest.approx(1 * vol * counts[1, i - 1] / (dv * 20 * 10 * 4))
        assert out[i - 1, 3] == pytest.approx(2 * vol * counts[2, i - 1] / (dv * 20 * 20 * 4))


def test_ideal_gas_g_is_one():
    rng = np.random.default_rng(9)
    L, n = 14.0, 3000
    pos, itype = _random_frame(rng, n, L, 1)
    ls = O.rdf_lsmax(3, [L] * 3, 5.0, 0.25)
    c = O.rdf_counts(pos, itype, 3, 1, [L] * 3, (1, 1, 1), 5.0, 0.25)
    g = O.rdf_gmix(3, 1, ls, 0.25, L ** 3, [n], 1, c)
    assert abs(g[4:, 1].mean() - 1.0) < 0.02


def test_first_minimum_parabola():
    """Counts shaped so that g has one peak then a parabolic dip at a known r."""
    nsp, lsmax, dr, vol, n = 1, 40, 0.1, 1.0, 1000
    r = np.arange(1, lsmax + 1) * dr
    dv = 4 * np.pi / 3 * ((r + dr / 2) ** 3 - (r - dr / 2) ** 3)
    rmin_true = 2.0 + 0.037
    gwant = np.where(r < 1.0, 3.0 * r, 3.0 - 2.0 * (r - 1.0)) * 0 + 1.0
    gwant = 1.0 + 0.5 * (r - rmin_true) ** 2
    gwant[r < 1.0] = 3.0 * r[r < 1.0]  # rises to a peak at r ~ 1
    gwant[(r >= 1.0) & (r < 1.05)] = 3.2
    counts = np.round(gwant * dv * n * n / (2 * vol) * 1e3).astype(np.int64)[None, :]
    rm = O.first_minimum(3, nsp, lsmax, dr, vol * 1e3, n, counts)
    assert abs(rm - rmin_true) < 0.02
