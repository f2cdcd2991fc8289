"""The project autotuner's cell-side and K-split choices (no counterpart in the
reference, whose launch shapes are fixed: rdf.cuf:208-214, sq.cuf:175-179).

Cell side: cells of at least f cutoffs for f = 0.5 (walked in x-rows), 1.25,
1.5 and 2 (fixed through TRJ_RDF_CELL_SCALE) and the autotuner's own pick, on
dense, sparse, slab, 2D, narrow and 1M-atom frames: counts bit-exact against
the oracle's cell lists, and the grid the library reports is the one the rule
gives.  K-split: the FP64 contraction with
the split length the autotuner picks equals the default split to 1e-12 per
shell, and its rho vectors agree with the oracle's long-double sums."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from trj_analysis_amd import trjhip as T

pytestmark = pytest.mark.gpu


def _frame(n, nsp, seed, density):
    rng = np.random.default_rng(seed)
    L = (n / density) ** (1 / 3)
    return rng.random((n, 3)) * L, rng.integers(1, nsp + 1, n).astype(np.int32), [L] * 3


def _cells(L, rc, f):
    return int(math.floor(L / (rc * f * (1.0 + 1e-6))))


@pytest.mark.parametrize("scale", [0.5, 1.25, 1.5, 2.0])
@pytest.mark.parametrize("density", [0.85, 0.08])
def test_fixed_cell_side_bit_exact(monkeypatch, scale, density):
    nsp, rc, dr = 3, 5.0, 0.01
    pos, it, box = _frame(40000, nsp, 11, density)
    monkeypatch.setenv("TRJ_RDF_CELL_SCALE", str(scale))
    print("case: cell side", scale, "cutoffs, density", density, flush=True)
    with T.Context(0) as ctx:
        st = ctx.rdf_init(T.rdf_params(3, nsp, dr, rc), box)
        ctx.frame(pos, it, box, what=T.TRJ_WANT_RDF)
        got = ctx.rdf_counts()
        grid = ctx.rdf_cell_grid()
        tu = ctx.rdf_tuning()
    print("grid", grid, "tuning", tu, flush=True)
    nc = _cells(box[0], rc, scale)
    assert grid["cells"] == [nc] * 3 and grid["side_over_cutoff"] == scale
    want = O.rdf_counts(pos, it, 3, nsp, box, (1, 1, 1), rc, dr, lsmax=st.lsmax, method="cells")
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("geom", ["slab", "2d", "rect"])
def test_half_cutoff_cells_geometries(monkeypatch, geom):
    """Cells of half the cutoff walked in x-rows (reach 2): a z-confined slab
    (non-periodic z, atoms up to 0.2 L outside the walls), a 2D frame and a
    rectangular box with 7 cells along x: counts bit-exact against the
    oracle's cell lists."""
    rng = np.random.default_rng(41)
    rc, dr, nsp = 4.0, 0.01, 3
    if geom == "2d":
        n, L = 20000, [160.0, 150.0, 1.0]
        pos = np.zeros((n, 3))
        pos[:, :2] = rng.random((n, 2)) * L[:2]
        per, ndim = (1, 1, 0), 2
    elif geom == "slab":
        n, L = 40000, [40.0, 44.0, 30.0]
        pos = rng.random((n, 3)) * L
        pos[:, 2] = rng.uniform(-0.2, 1.2, n) * L[2]
        per, ndim = (1, 1, 0), 3
    else:
        n, L = 30000, [14.5, 60.0, 40.0]  # 7 half-cutoff cells along x
        pos = rng.random((n, 3)) * L
        per, ndim = (1, 1, 1), 3
    it = rng.integers(1, nsp + 1, n).astype(np.int32)
    monkeypatch.setenv("TRJ_RDF_CELL_SCALE", "0.5")
    print("case: half-cutoff cells", geom, flush=True)
    with T.Context(0) as ctx:
        st = ctx.rdf_init(T.rdf_params(ndim, nsp, dr, rc, per), L)
        ctx.frame(np.ascontiguousarray(pos[:, :ndim]), it, L, what=T.TRJ_WANT_RDF)
        got = ctx.rdf_counts()
        grid = ctx.rdf_cell_grid()
    print("grid", grid, flush=True)
    assert grid["side_over_cutoff"] == 0.5
    want = O.rdf_counts(pos, it, ndim, nsp, L, per, rc, dr, lsmax=st.lsmax, method="cells")
    np.testing.assert_array_equal(got, want)


def test_half_cutoff_cells_1m():
    """uniform_1m (1M atoms, 3 species, rc = 5) on cells of half the cutoff:
    counts bit-exact against the oracle's cell lists."""
    from trj_analysis_amd import workloads as W
    f = W.uniform_1m(seed=5)
    wl = W.UNIFORM_1M
    import os
    os.environ["TRJ_RDF_CELL_SCALE"] = "0.5"
    try:
        print("case: half-cutoff cells uniform_1m", flush=True)
        with T.Context(0) as ctx:
            st = ctx.rdf_init(T.rdf_params(3, wl.nsp, wl.deltar, wl.rcrdf), f.box)
            ctx.frame(f.pos, f.itype, f.box, what=T.TRJ_WANT_RDF)
            got = ctx.rdf_counts()
            grid = ctx.rdf_cell_grid()
    finally:
        del os.environ["TRJ_RDF_CELL_SCALE"]
    print("grid", grid, flush=True)
    assert grid["side_over_cutoff"] == 0.5
    want = O.rdf_counts(f.pos, f.itype, 3, wl.nsp, f.box, (1, 1, 1), wl.rcrdf, wl.deltar, lsmax=st.lsmax,
                        method="cells")
    np.testing.assert_array_equal(got, want)


def test_autotuned_cell_side_sparse():
    """A sparse fluid (about 7 atoms per cell of one cutoff): the tuner times
    the three cell sides; whichever it keeps, two frames count bit-exact."""
    nsp, rc, dr = 2, 5.0, 0.01
    frames = [_frame(30000, nsp, s, 0.06) for s in (21, 22)]
    box = frames[0][2]
    print("case: autotuned cell side, sparse", flush=True)
    with T.Context(0) as ctx:
        st = ctx.rdf_init(T.rdf_params(3, nsp, dr, rc), box)
        for pos, it, _ in frames:
            ctx.frame(pos, it, box, what=T.TRJ_WANT_RDF)
        got = ctx.rdf_counts()
        grid = ctx.rdf_cell_grid()
        tu = ctx.rdf_tuning()
    print("grid", grid, "tuning", tu, flush=True)
    assert grid["side_over_cutoff"] in (0.5, 1.0, 1.25, 1.5)
    assert grid["cells"][0] == _cells(box[0], rc, grid["side_over_cutoff"])
    assert tu["tried"] >= 4
    want = sum(O.rdf_counts(pos, it, 3, nsp, box, (1, 1, 1), rc, dr, lsmax=st.lsmax, method="cells")
               for pos, it, _ in frames)
    np.testing.assert_array_equal(got, want)


def test_fp64_split_tuner():
    """FP64 contraction on the 128^3 half-space grid (512 tiles) for 100k
    atoms: split lengths 16384 and 32768 give different split counts, the
    tuner times them in both workgroup orders; shells equal the untuned
    context's to 1e-12 and ten rho vectors match the oracle to 1e-10 relative
    to sqrt(N_s)."""
    nsp = 3
    pos, it, box = _frame(100_000, nsp, 31, 0.85)
    bsc = [1.0, 0.6, -0.3]
    out = {}
    for tuned in (True, False):
        flags = 0 if tuned else T.TRJ_SQ_NO_AUTOTUNE
        p = T.sq_params(3, nsp, 1e9, bsc=bsc, qset=T.TRJ_QSET_HALFSPACE, grid=(128, 128, 128),
                        precision=T.TRJ_PREC_FP64, flags=flags)
        print("case: fp64 split tuner, tuned", tuned, flush=True)
        with T.Context(0) as ctx:
            ctx.sq_init(p, box)
            ctx.frame(pos, it, box, what=T.TRJ_WANT_SQ)
            sqf, sqfp, nq = ctx.sq_shells()
            tu = ctx.sq_tuning()
            kv = np.array([[1, 0, 0], [0, 1, 0], [3, -2, 5], [-7, 4, 1], [12, 0, 9], [-20, 31, 2], [63, -64, 63],
                           [5, 5, 5], [-1, -1, 1], [40, -33, 17]], dtype=np.int32)
            rho = ctx.sq_rho(kv) if tuned else None
        print("tuning", tu, flush=True)
        out[tuned] = (sqf, sqfp, nq, tu, rho)
    a, b = out[True], out[False]
    assert a[3]["tried"] == 4 and a[3]["atoms_per_split"] in (16384, 32768) and a[3]["xcd_grouped"] in (0, 1)
    assert b[3]["tried"] == 0 and b[3]["atoms_per_split"] == 0 and b[3]["xcd_grouped"] == 0
    np.testing.assert_array_equal(a[2], b[2])
    m = b[0] > 0
    assert np.max(np.abs(a[0][m] - b[0][m]) / b[0][m]) <= 1e-12
    orho = O.sq_rho(O.wrap(pos, box, (1, 1, 1)), it, nsp, box, kv)
    ns = np.bincount(it, minlength=nsp + 1)[1:]
    err = np.abs(a[4] - orho) / np.sqrt(ns)[None, :]
    print("rho vs oracle: max |drho|/sqrt(N_s) %.3e" % err.max(), flush=True)
    assert err.max() <= 1e-10
