"""Pins of the oracle's number-fluctuation sums (s2D3_sh): closed forms and
an independent integer-vector count on a simple-cubic lattice."""
import itertools

import numpy as np

from oracle import oracle as O


def test_single_atom_known_distance():
    L = 20.0
    pos = np.array([[10.0, 10.0, 10.0]])
    pts = np.array([[13.0 / L, 14.0 / L, 10.0 / L]])     # distance 5
    dr = 0.1
    gi, g2i = O.s2n_sums(pts, pos, 3, [L] * 3, (1, 1, 1), 8.0, dr)
    k = int(np.floor(5.0 / dr + 0.5))
    want = (np.arange(len(gi)) >= k).astype(np.int64)
    np.testing.assert_array_equal(gi, want)
    np.testing.assert_array_equal(g2i, want)


def test_lattice_site_counts_integer_vectors():
    a, m = 1.0, 12
    L = a * m
    g = np.arange(m) * a
    pos = np.array(list(itertools.product(g, g, g)))
    pts = np.array([[3.0 / L, 5.0 / L, 7.0 / L]])         # on a lattice site
    rc, dr = 3.2, 0.05
    gi, g2i = O.s2n_sums(pts, pos, 3, [L] * 3, (1, 1, 1), rc, dr)
    # N_k: integer vectors v with |v| < rc and nint(|v|/dr) <= k
    rng = range(-4, 5)
    rr = np.array([np.sqrt(i * i + j * j + k * k) for i in rng for j in rng for k in rng])
    rr = rr[rr < rc]
    bins = np.floor(rr / dr + 0.5).astype(int)
    want = np.array([np.sum(bins <= k) for k in range(len(gi))])
    np.testing.assert_array_equal(gi, want)
    np.testing.assert_array_equal(g2i, want * want)


def test_total_at_lsmax_is_atoms_within_cutoff():
    rng = np.random.default_rng(5)
    L = 15.0
    pos = rng.random((500, 3)) * L
    pts = rng.random((20, 3))
    rc, dr = 4.0, 0.05
    gi, g2i = O.s2n_sums(pts, pos, 3, [L] * 3, (1, 1, 1), rc, dr)
    tot = 0
    tot2 = 0
    for p in pts:
        d = p * L - pos
        d -= L * np.round(d / L)
        c = int(np.sum((d ** 2).sum(1) < rc * rc))
        tot += c
        tot2 += c * c
    assert gi[-1] == tot and g2i[-1] == tot2
