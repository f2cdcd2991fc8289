"""GPU parity of the confined-slice analysis (twoDstruc_3D: rdfxy, sqf3xy /
sqf2xy) against the oracle, through the C-ABI: raw slice pair counts and
slice counts bit-exact, the per-frame normalised accumulators bit-exact, the
S(Q) slice sums to the stated tolerance."""
import numpy as np
import pytest

from oracle import oracle as O
from trj_analysis_amd import trjhip as T

pytestmark = pytest.mark.gpu


def _frames(nf, n, L, zlo, zhi, seed, nsp=2):
    rng = np.random.default_rng(seed)
    it = rng.integers(1, nsp + 1, n).astype(np.int32)
    out = []
    for _ in range(nf):
        pos = np.empty((n, 3))
        pos[:, :2] = rng.random((n, 2)) * L
        pos[:, 2] = zlo + rng.random(n) * (zhi - zlo)
        out.append(pos)
    return out, it


def _run_rdf(frames, it, box, rc, dr, zs, zg, nsp=2, q=None):
    with T.Context(0) as ctx:
        ctx.set_slices(zs, zg)
        flags = T.TRJ_RDF_CHARGE if q is not None else 0
        st = ctx.rdf_init(T.rdf_params(3, nsp, dr, rc, (1, 1, 0), flags=flags), box)
        for k, pos in enumerate(frames):
            ctx.frame(pos, it, box, what=T.TRJ_WANT_RDF, charge=None if q is None else q[k])
        h, c = ctx.rdf_xy()
        cnt, cntp = ctx.slice_counts()
        g = ctx.rdf_xy_gmix()
        gq = ctx.rdf_xy_qq() if q is not None else None
        gcc = ctx.rdf_xy_gcc() if q is not None else None
        with pytest.raises(T.TrjError):
            ctx.rdf_counts()                           # unsliced results are refused
        return st.lsmax, h, c, cnt, cntp, g, gq, gcc


@pytest.mark.parametrize("case", ["cells_shift", "cells_noshift", "allpairs"])
def test_rdf_xy_bit_exact(case):
    L, rc = {"cells_shift": (32.0, 6.0), "cells_noshift": (20.0, 5.0), "allpairs": (11.0, 5.0)}[case]
    n = {"cells_shift": 6000, "cells_noshift": 3000, "allpairs": 800}[case]
    frames, it = _frames(2, n, L, -10.0, 22.0, seed=len(case))
    box = [L, L, 40.0]
    zs, zg, dr = [-5.0, 5.0, 15.0], 6.0, 0.05
    print("case: rdf_xy", case, flush=True)
    ls, h, c, cnt, cntp, g, _, _ = _run_rdf(frames, it, box, rc, dr, zs, zg)
    oc_tot = None
    oh = None
    for pos in frames:
        oc, ocnt, ocntp, _, _ = O.rdf_xy_counts(pos, it, 2, box, rc, dr, zs, zg, lsmax=ls)
        oc_tot = oc if oc_tot is None else oc_tot + oc
        oh = O.rdf_xy_accumulate(2, oc, ocntp, oh)
    np.testing.assert_array_equal(c, oc_tot)
    np.testing.assert_array_equal(cnt, ocnt)
    np.testing.assert_array_equal(cntp, ocntp)
    h_pairs = np.stack([oh[:, j, l] for j in range(2) for l in range(j, 2)], axis=1)
    np.testing.assert_array_equal(h, h_pairs)
    og = O.rdf_xy_gmix(2, ls, dr, box, zg, 2, oh)
    np.testing.assert_allclose(g, og, rtol=1e-14, atol=0)
    assert c.sum() > 10000 or case == "allpairs"


def test_rdf_xy_charges():
    frames, it = _frames(2, 3000, 24.0, -8.0, 18.0, seed=7)
    box = [24.0, 24.0, 34.0]
    zs, zg, rc, dr = [0.0, 10.0], 8.0, 5.0, 0.05
    rng = np.random.default_rng(8)
    qs = [np.where(it == 1, 0.9, -0.9) + 0.05 * rng.standard_normal(len(it)) for _ in frames]
    print("case: rdf_xy charges", flush=True)
    ls, h, c, cnt, cntp, g, gq, gcc = _run_rdf(frames, it, box, rc, dr, zs, zg, q=qs)
    acc = np.zeros((2, ls))
    for pos, q in zip(frames, qs):
        _, ocnt, _, fixed, scale = O.rdf_xy_counts(pos, it, 2, box, rc, dr, zs, zg, lsmax=ls, q=q)
        sl = np.array([O.slice_of(z, zs, zg) for z in pos[:, 2]])
        for s in range(2):
            q2 = np.mean(q[sl == s] ** 2)
            acc[s] += (fixed[s] / scale) / (float(ocnt[s]) ** 2 * q2)
    # the slice q^2 sums are FP64 atomics (last-bit order dependence); bins
    # where two frames cancel need an absolute bound
    np.testing.assert_allclose(gq, acc, rtol=1e-12, atol=1e-12 * np.abs(acc).max())
    vol = np.prod(box)
    i = np.arange(1, ls - 1)
    ri = i * dr
    zg64 = float(np.float32(zg))                       # the FP32 zgrid, in FP64 arithmetic
    dv = np.pi * ((ri + dr / 2) ** 2 - (ri - dr / 2) ** 2) * zg64
    want = vol * (zg64 / box[2]) * acc[:, : ls - 2].T / (dv[:, None] * 2)
    np.testing.assert_allclose(gcc, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    m = np.abs(b) > 1e-300
    return float(np.max(np.abs(a[m] - b[m]) / np.abs(b[m]))) if m.any() else 0.0


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp16x3", 1e-5)])
def test_sq_xy(precision, tol):
    frames, it = _frames(2, 4000, 26.0, -6.0, 14.0, seed=11)
    box = [26.0, 26.0, 28.0]
    zs, zg = [-2.0, 9.0], 7.0
    bsc = [1.0, 0.7]
    rng = np.random.default_rng(12)
    qs = [np.where(it == 1, 0.6, -0.6) + 0.05 * rng.standard_normal(len(it)) for _ in frames]
    print("case: sq_xy", precision, flush=True)
    with T.Context(0) as ctx:
        ctx.set_slices(zs, zg)
        info = ctx.sq_init(T.sq_params(3, 2, 8.0, 5.0, bsc, precision=precision, flags=T.TRJ_SQ_CHARGE), box)
        for pos, q in zip(frames, qs):
            ctx.frame(pos, it, box, what=T.TRJ_WANT_SQ, charge=q)
        f, fp, fq = ctx.sq_xy(charges=True)
        cols, pcols, qcols = ctx.sq_xy_columns(charges=True)
        with pytest.raises(T.TrjError):
            ctx.sq_shells()
    g = O.sq_setup(2, box, 8.0, 5.0)
    assert info.nqmax == g.nqmax and info.nqmin == g.nqmin
    acc = None
    for pos, q in zip(frames, qs):
        acc = O.sq_frame_xy(pos, it, 2, box, g, bsc, zs, zg, q=q, acc=acc)
    of, ofp, ofq = acc
    print("rel", _rel(f, of), _rel(fp, ofp), _rel(fq, ofq), flush=True)
    assert _rel(f, of) <= tol and _rel(fp, ofp) <= tol and _rel(fq, ofq) <= tol
    nq = O.sq_nq(2, g)
    i = np.arange(1, g.nqmax + 1)
    den = 2.0 * nq[1:]
    assert np.allclose(cols[:, 0], (i.astype(np.float32) * np.float32(g.dq)).astype(np.float64))
    np.testing.assert_allclose(cols[:, 1:], (of[:, 1:] / den).T, rtol=max(tol, 1e-12))
    np.testing.assert_allclose(pcols[1, :, 1:], (ofp[:, 1, 1:] / den).T, rtol=max(tol, 1e-12))
    sl = np.array([O.slice_of(z, zs, zg) for z in frames[-1][:, 2]])
    q2 = np.array([np.mean(qs[-1][sl == s] ** 2) for s in range(2)])   # last frame (sq.cuf:135-138)
    np.testing.assert_allclose(qcols[:, 1:], (ofq[:, 1:] / q2[:, None] / den).T, rtol=max(tol, 1e-12))


def test_slice_argument_checks():
    print("case: slice argument checks", flush=True)
    with T.Context(0) as ctx:
        with pytest.raises(T.TrjError):
            ctx.set_slices([0.0] * 17, 1.0)
        with pytest.raises(T.TrjError):
            ctx.set_slices([0.0], 0.0)
        ctx.set_slices([0.0], 2.0)
        with pytest.raises(T.TrjError):                # x, y must be periodic
            ctx.rdf_init(T.rdf_params(3, 1, 0.1, 3.0, (1, 0, 0)), [10.0, 10.0, 10.0])
        ctx.rdf_init(T.rdf_params(3, 1, 0.1, 3.0, (1, 1, 0)), [10.0, 10.0, 10.0])
        with pytest.raises(T.TrjError):                # after init
            ctx.set_slices([1.0], 2.0)


def test_sq_xy_species_charges():
    """Charges constant per species: the slice charge sums are formed from the
    (slice, species) channels, no charge channel; S_xy_qq matches the oracle."""
    frames, it = _frames(2, 4000, 26.0, -6.0, 14.0, seed=13)
    box = [26.0, 26.0, 28.0]
    zs, zg = [-2.0, 9.0], 7.0
    bsc = [1.0, 0.7]
    q = np.where(it == 1, 0.6, -0.6)
    print("case: sq_xy species charges", flush=True)
    with T.Context(0) as ctx:
        ctx.set_slices(zs, zg)
        ctx.sq_init(T.sq_params(3, 2, 8.0, 5.0, bsc, flags=T.TRJ_SQ_CHARGE), box)
        for pos in frames:
            ctx.frame(pos, it, box, what=T.TRJ_WANT_SQ, charge=q)
        f, fp, fq = ctx.sq_xy(charges=True)
    g = O.sq_setup(2, box, 8.0, 5.0)
    acc = None
    for pos in frames:
        acc = O.sq_frame_xy(pos, it, 2, box, g, bsc, zs, zg, q=q, acc=acc)
    of, ofp, ofq = acc
    assert _rel(f, of) <= 1e-12 and _rel(fp, ofp) <= 1e-12 and _rel(fq, ofq) <= 1e-12
    assert np.abs(ofq).max() > 0
