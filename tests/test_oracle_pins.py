"""Pins of the oracle's output normalisation and q-vector member rule by what
the mathematics fixes, independently of the oracle's own formulas:

  * or_sq_member: exact integer lattice-point counts per shell (away from
    the FP32 edges), on a cubic and a 1:2:1 box;
  * or_sq_columns (printSQ, log.f90:317-401):
      - perfect lattices: Bragg sums are exact, so S_NN, S_jj and the
        Bhatia-Thornton S_cc of an ordered binary alloy have closed forms;
      - ideal gas: S_jj -> 1, S_NN -> <b^2>/<b>^2; an ideal binary mixture
        with unit scattering lengths has s11 -> x1, s22 -> x2, s12 -> 0,
        S_cc -> x1 x2;
      - exact invariances: the same frame accumulated k times with Nconf = k,
        and swapping the species labels;
  * or_rdf_xy_gmix (log.f90:494-514): an ideal slab gives g_xy -> the value
    the reference's centred cylindrical shell fixes for its floor bins.

The shell counts nq used for normalisation are the vectors actually summed
(the single-atom frame: one |rho|^2 = 1 per summed vector), so the checks do
not depend on the reference's host count at the boundary shell."""
import math

import numpy as np
import pytest

from oracle import oracle as O


# ---------------------------------------------------------------- or_sq_member

def _member_counts(g, ndim=3):
    """Per-shell counts of or_sq_member over the kernel's octant box."""
    out = {}
    kz = g.kext[2] if ndim == 3 else 1
    for a in range(g.kext[0]):
        for b in range(g.kext[1]):
            for c in range(kz):
                s = O.sq_member(ndim, g, a, b, c)
                if s >= 0:
                    out.setdefault((a, b, c), s)
    return out


@pytest.mark.parametrize("box,scale,qmax,qmin", [
    ([13.0, 13.0, 13.0], (1, 1, 1), 9.0, 9.0),     # l = 18 -> 16 after the 4-adjust
    ([9.5, 9.5, 9.5], (1, 1, 1), 10.0, 8.0),
    ([10.0, 20.0, 10.0], (2, 1, 2), 8.0, 8.0),     # fk_x = fk_z = 2 fk_y exactly in FP32
])
def test_sq_member_integer_lattice_counts(box, scale, qmax, qmin):
    """In units of dq = min fk, q^2/dq^2 = m = sum (scale_a n_a)^2 is an integer
    and qmin'/dq = min_a scale_a l_a; the member rule is 0 < m <= (qmin'/dq)^2
    and the shell floor(sqrt(m) + 0.501) if <= nqmax.  Vectors on the sphere
    (m = (qmin'/dq)^2) or within 1e-4 of a shell edge are left out: there the
    FP32 rounding decides, which the golden n(Q) columns pin instead."""
    g = O.sq_setup(3, box, qmax, qmin)
    lmax = min(s * l for s, l in zip(scale, g.l))
    got = _member_counts(g)
    want, have = {}, {}
    for a in range(g.kext[0]):
        for b in range(g.kext[1]):
            for c in range(g.kext[2]):
                m = (scale[0] * a) ** 2 + (scale[1] * b) ** 2 + (scale[2] * c) ** 2
                if m == 0 or m == lmax * lmax:
                    continue
                t = math.sqrt(m) + 0.501
                if abs(t - round(t)) < 1e-4:
                    continue
                s = math.floor(t)
                if m < lmax * lmax and s <= g.nqmax:
                    want[s] = want.get(s, 0) + 1
                if (a, b, c) in got:
                    have[got[(a, b, c)]] = have.get(got[(a, b, c)], 0) + 1
    assert sum(want.values()) > 500
    assert have == want


# ---------------------------------------------------------------- or_sq_columns

def _summed_counts(ndim, box, g):
    """Vectors summed per shell: a single atom adds |rho|^2 = 1 per vector."""
    sqf, _ = O.sq_frame(np.array([[0.37, 0.11, 0.73]]), [1], ndim, 1, box, g, [1.0])
    return np.rint(sqf).astype(np.int64)


def _bragg_counts(g, m, residue):
    """Per-shell count of summed vectors with every component = residue mod m
    (the octant set, then the axis vectors of shells nqmin+1..nqmax)."""
    out = np.zeros(g.nqmax + 1, dtype=np.int64)
    for (a, b, c), s in _member_counts(g).items():
        if a % m == residue and b % m == residue and c % m == residue:
            out[s] += 1
    if residue == 0:
        for k in range(max(1, g.nqmin + 1), g.nqmax + 1):
            if k % m == 0:
                out[k] += 3
    return out


def _lattice(m, a, shift):
    idx = np.arange(m)
    ijk = np.stack(np.meshgrid(idx, idx, idx, indexing="ij"), -1).reshape(-1, 3)
    return ijk, (ijk + 0.25) * a + shift


def test_sq_columns_lattice_bragg_exact():
    """A perfect simple-cubic lattice, m = 6 sites a side: |rho(q)|^2 = N^2 on
    the reciprocal lattice (every n = 0 mod 6) and 0 elsewhere.  Over Nconf
    translated copies, S_NN(shell) = N * nB0 / nq exactly.  With the species
    alternating (NaCl order, i + j + k even or odd) each species is an fcc
    lattice of N/2 sites: rho_1 = rho_2 = N/2 on nB0 vectors, rho_1 = -rho_2 =
    N/2 on the superlattice vectors (every n = 3 mod 6, nB3), so
      S_jj(mix) = (N/2)(nB0 + nB3)/nq,   S_NN = N nB0/nq,
      S_cc = (N/4) nB3/nq  (Bhatia-Thornton: x1 x2 N on the superlattice)."""
    m, a = 6, 1.1
    L = m * a
    box = [L] * 3
    g = O.sq_setup(3, box, 12.0, 12.0)
    nq = _summed_counts(3, box, g)
    nb0, nb3 = _bragg_counts(g, m, 0), _bragg_counts(g, m, 3)
    assert nb0[1:].sum() >= 6 and nb3[1:].sum() >= 4
    ijk, _ = _lattice(m, a, 0.0)
    n = len(ijk)
    rng = np.random.default_rng(4)
    nconf = 3
    # one species, b = 1
    sqf = sqfp = None
    for _ in range(nconf):
        pos = _lattice(m, a, rng.random(3) * a)[1] % L
        sqf, sqfp = O.sq_frame(pos, np.ones(n, np.int32), 3, 1, box, g, [1.0], sqf, sqfp)
    cols, mix = O.sq_columns(1, g, n, [n], [1.0], nconf, sqf, sqfp, nq)
    sh = np.arange(1, g.nqmax + 1)
    ok = nq[sh] > 0
    want = n * nb0[sh] / np.maximum(nq[sh], 1)
    np.testing.assert_allclose(cols[ok, 1], want[ok], rtol=1e-12, atol=1e-18)
    np.testing.assert_allclose(mix[ok, 1], want[ok], rtol=1e-12, atol=1e-18)
    np.testing.assert_array_equal(cols[ok, 2], nq[sh][ok])
    np.testing.assert_allclose(cols[:, 0], sh * np.float32(g.dq).astype(np.float64), rtol=1e-7)
    # ordered binary alloy, unit scattering lengths (the 7-column binary table)
    it = np.where(ijk.sum(1) % 2 == 0, 1, 2).astype(np.int32)
    sqf = sqfp = None
    for _ in range(nconf):
        pos = _lattice(m, a, rng.random(3) * a)[1] % L
        sqf, sqfp = O.sq_frame(pos, it, 3, 2, box, g, [1.0, 1.0], sqf, sqfp)
    cols, mix = O.sq_columns(2, g, n, [n // 2, n // 2], [1.0, 1.0], nconf, sqf, sqfp, nq)
    assert cols.shape[1] == 7
    q = np.maximum(nq[sh], 1)
    np.testing.assert_allclose(cols[ok, 1], (n * nb0[sh] / q)[ok], rtol=1e-12, atol=1e-18)
    # S_cc subtracts terms of order N: rounding leaves about 1e-16 N
    np.testing.assert_allclose(cols[ok, 2], (n / 4 * nb3[sh] / q)[ok], rtol=1e-12, atol=1e-14 * n)
    for s in (1, 2):
        np.testing.assert_allclose(mix[ok, s], (n / 2 * (nb0[sh] + nb3[sh]) / q)[ok], rtol=1e-12, atol=1e-18)
    # equal scattering lengths other than 1: sqf = b^2 |rho|^2 and <b>^2 = b^2
    sqf2 = None
    for _ in range(nconf):
        pos = _lattice(m, a, rng.random(3) * a)[1] % L
        sqf2, _ = O.sq_frame(pos, it, 3, 2, box, g, [2.5, 2.5], sqf2, None)
    cols2, _ = O.sq_columns(2, g, n, [n // 2, n // 2], [2.5, 2.5], nconf, sqf2, sqfp, nq)
    assert cols2.shape[1] == 3
    np.testing.assert_allclose(cols2[ok, 1], (n * nb0[sh] / q)[ok], rtol=1e-12, atol=1e-18)


def _ideal_gas(n, L, x1, nconf, seed):
    rng = np.random.default_rng(seed)
    n1 = int(round(x1 * n))
    it = np.concatenate([np.ones(n1, np.int32), np.full(n - n1, 2, np.int32)])
    rng.shuffle(it)
    return [rng.random((n, 3)) * L for _ in range(nconf)], it, n1


def test_sq_columns_ideal_gas_limits():
    """Uniform independent positions: E|rho_j(q)|^2 = N_j on every reciprocal
    vector q != 0, so S_jj -> 1 and S_NN -> <b^2>/<b>^2; with unit scattering
    lengths s11 -> x1, s22 -> x2, s12 -> 0 and S_cc -> x1 x2.  Each |rho|^2/N is
    close to Exp(1), so a mean over V vectors has relative spread 1/sqrt(V):
    the pooled means are held to 5 sigma and each shell to 6 sigma."""
    n, L, x1, nconf = 3000, 15.5, 0.3, 4
    frames, it, n1 = _ideal_gas(n, L, x1, nconf, seed=11)
    box = [L] * 3
    g = O.sq_setup(3, box, 8.0, 8.0)
    nq = _summed_counts(3, box, g)
    ntype = [n1, n - n1]
    x = np.array(ntype) / n
    bsc = [1.0, 0.5]
    acc = {}
    for b in (bsc, [1.0, 1.0]):
        sqf = sqfp = None
        for f in frames:
            sqf, sqfp = O.sq_frame(f, it, 3, 2, box, g, b, sqf, sqfp)
        acc[tuple(b)] = O.sq_columns(2, g, n, ntype, b, nconf, sqf, sqfp, nq)
    sh = np.arange(1, g.nqmax + 1)
    w = nq[sh] * nconf
    ok = w >= 200
    V = w[ok].sum()
    assert V > 10000
    sig = 1.0 / math.sqrt(V)

    def pooled(col):
        return float(np.sum(col[ok] * w[ok]) / V)

    cols, mix = acc[tuple(bsc)]
    b2 = float(np.dot(x, np.square(bsc)) / np.dot(x, bsc) ** 2)   # <b^2>/<b>^2 = 1.1243
    assert abs(pooled(cols[:, 1]) / b2 - 1.0) < 5 * sig
    for s in (1, 2):
        assert abs(pooled(mix[:, s]) - 1.0) < 5 * sig
        assert np.all(np.abs(mix[ok, s] - 1.0) < 6.0 / np.sqrt(w[ok]))
    cols, mix = acc[(1.0, 1.0)]
    assert cols.shape[1] == 7
    snn, scc, s11, s22, s12 = (cols[:, k] for k in (1, 2, 3, 4, 5))
    assert abs(pooled(snn) - 1.0) < 5 * sig
    assert abs(pooled(s11) - x[0]) < 5 * sig * x[0]
    assert abs(pooled(s22) - x[1]) < 5 * sig * x[1]
    assert abs(pooled(s12)) < 5 * sig
    assert abs(pooled(scc) - x[0] * x[1]) < 5 * sig * x[0] * x[1] * 3
    np.testing.assert_array_equal(cols[:, 6], nq[sh])


def test_sq_columns_exact_invariances():
    """The same frame accumulated k times with Nconf = k gives the columns of
    one frame with Nconf = 1; swapping the species labels (and their
    scattering lengths) swaps s11 <-> s22 and the two mix columns and leaves
    S_NN, S_cc and s12 unchanged."""
    n, L = 800, 10.0
    frames, it, n1 = _ideal_gas(n, L, 0.35, 1, seed=3)
    box = [L] * 3
    g = O.sq_setup(3, box, 7.0, 5.0)
    nq = _summed_counts(3, box, g)
    ntype = [n1, n - n1]
    one = O.sq_frame(frames[0], it, 3, 2, box, g, [1.0, 1.0])
    c1, m1 = O.sq_columns(2, g, n, ntype, [1.0, 1.0], 1, *one, nq)
    sqf = sqfp = None
    for _ in range(5):
        sqf, sqfp = O.sq_frame(frames[0], it, 3, 2, box, g, [1.0, 1.0], sqf, sqfp)
    c5, m5 = O.sq_columns(2, g, n, ntype, [1.0, 1.0], 5, sqf, sqfp, nq)
    np.testing.assert_allclose(c5, c1, rtol=1e-13)
    np.testing.assert_allclose(m5, m1, rtol=1e-13)
    sw = (3 - it).astype(np.int32)
    two = O.sq_frame(frames[0], sw, 3, 2, box, g, [1.0, 1.0])
    c2, m2 = O.sq_columns(2, g, n, ntype[::-1], [1.0, 1.0], 1, *two, nq)
    np.testing.assert_allclose(c2[:, [0, 1, 2, 5, 6]], c1[:, [0, 1, 2, 5, 6]], rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(c2[:, 3], c1[:, 4], rtol=1e-12)
    np.testing.assert_allclose(c2[:, 4], c1[:, 3], rtol=1e-12)
    np.testing.assert_allclose(m2[:, 1], m1[:, 2], rtol=1e-12)
    np.testing.assert_allclose(m2[:, 2], m1[:, 1], rtol=1e-12)
    # weighted: the same relabelling with b = (1, 0.4) -> (0.4, 1)
    w1 = O.sq_frame(frames[0], it, 3, 2, box, g, [1.0, 0.4])
    w2 = O.sq_frame(frames[0], sw, 3, 2, box, g, [0.4, 1.0])
    a, _ = O.sq_columns(2, g, n, ntype, [1.0, 0.4], 1, *w1, nq)
    b, _ = O.sq_columns(2, g, n, ntype[::-1], [0.4, 1.0], 1, *w2, nq)
    np.testing.assert_allclose(a, b, rtol=1e-12)


# ---------------------------------------------------------------- or_rdf_xy_gmix

def test_rdf_xy_gmix_ideal_slab():
    """Uniform atoms in a slab box Lx = Ly = 30, Lz = 20, two slices of
    thickness zgrid = 4: the pair count of floor bin i, r in [(i-1)dr, i dr),
    is N_j N_l (2i-1) pi dr^2 / A for j < l and N_j (N_j - 1)/2 of that for
    j = l.  printrdf divides by the centred cylindrical shell
    pi ((r+dr/2)^2 - (r-dr/2)^2) zgrid = 2 pi i dr^2 zgrid and multiplies by
    (j/l + 1) V (zgrid/Lz), so an ideal slab gives
        g_jl(i) -> (2i - 1)/(2i) * c_jl,
    c_jl = N_j N_l / (countp_j countp_l), and N_j (N_j - 1)/countp_j^2 on the
    diagonal (countp: the slice's species counts of the reference's rule)."""
    Lx, Lz, zgrid, dr, rc = 30.0, 20.0, 4.0, 0.1, 4.5
    zs = [5.0, 15.0]
    nsp, n, nconf = 2, 20000, 3
    rng = np.random.default_rng(8)
    box = [Lx, Lx, Lz]
    hxy = None
    nsum = np.zeros((2, nsp))
    lsmax = None
    for _ in range(nconf):
        pos = rng.random((n, 3)) * box
        it = np.where(rng.random(n) < 0.3, 1, 2).astype(np.int32)
        counts, count, countp, _, _ = O.rdf_xy_counts(pos, it, nsp, box, rc, dr, zs, zgrid)
        lsmax = counts.shape[2]
        hxy = O.rdf_xy_accumulate(nsp, counts, countp, hxy)
        sl = np.array([O.slice_of(z, zs, zgrid) for z in pos[:, 2]])
        for s in range(2):
            for t in range(nsp):
                nsum[s, t] += np.sum((sl == s) & (it == t + 1))
    out = O.rdf_xy_gmix(nsp, lsmax, dr, box, zgrid, nconf, hxy)
    i = np.arange(1, lsmax - 1)
    shape = (2 * i - 1) / (2 * i)
    nj = nsum / nconf                                # mean slice species counts
    sel = i >= 10
    pair = 0
    for j in range(nsp):
        for l in range(j, nsp):
            for s in range(2):
                g = out[pair, :, s]
                # countp differs from the slice count by at most one atom
                c = (1.0 - 1.0 / nj[s, j]) if j == l else 1.0
                ratio = g[sel] / (shape[sel] * c)
                # pairs in the pooled bins, for the statistical spread
                npairs = (nj[s, j] * (nj[s, l] if j != l else nj[s, j] / 2) * nconf *
                          np.sum((2 * i[sel] - 1) * math.pi * dr * dr) / (Lx * Lx))
                tol = 5.0 / math.sqrt(npairs) + 3.0 / nj[s, j]
                assert abs(ratio.mean() - 1.0) < tol, (j, l, s, ratio.mean(), tol)
                per_bin = (nj[s, j] * (nj[s, l] if j != l else nj[s, j] / 2) * nconf *
                           (2 * i[sel] - 1) * math.pi * dr * dr / (Lx * Lx))
                assert np.all(np.abs(ratio - 1.0) < 6.0 / np.sqrt(per_bin) + 3.0 / nj[s, j])
            pair += 1
