"""GPU parity of the RDF path: counts bit-exact against the FP64 oracle."""
import numpy as np
import pytest

from oracle import oracle as O
from trj_analysis_amd import trjhip as T
from trj_analysis_amd import workloads as W

pytestmark = pytest.mark.gpu


def _gpu_counts(frames, ndim, nsp, deltar, rcrdf, periodic=(1, 1, 1)):
    with T.Context(0) as ctx:
        ctx.rdf_init(T.rdf_params(ndim, nsp, deltar, rcrdf, periodic), frames[0][2])
        for pos, it, box in frames:
            ctx.frame(pos, it, box, what=T.TRJ_WANT_RDF)
        return ctx.rdf_counts(), ctx


def _oracle_counts(frames, ndim, nsp, deltar, rcrdf, periodic=(1, 1, 1), method="brute"):
    ls = O.rdf_lsmax(ndim, frames[0][2], rcrdf, deltar)
    tot = None
    for pos, it, box in frames:
        c = O.rdf_counts(pos, it, ndim, nsp, box, periodic, rcrdf, deltar, lsmax=ls, method=method)
        tot = c if tot is None else tot + c
    return tot


def test_lj_binary_frames_bit_exact():
    fr = [(f.pos, f.itype, f.box) for f in W.lj_binary_4k(nframes=3)]
    print("case: lj_binary_4k x3", flush=True)
    g, _ = _gpu_counts(fr, 3, 2, 0.02, 4.0)
    o = _oracle_counts(fr, 3, 2, 0.02, 4.0)
    assert g.shape == o.shape == (3, 201)
    np.testing.assert_array_equal(g, o)
    assert g.sum() > 1_000_000


@pytest.mark.parametrize("layout", [T.TRJ_AOS, T.TRJ_SOA])
def test_layouts_and_dtypes(layout):
    rng = np.random.default_rng(4)
    L = 13.0
    pos = rng.random((2500, 3)) * L - 0.5 * L  # needs wrapping
    it = rng.integers(1, 4, 2500).astype(np.int32)
    print("case: layouts", layout, flush=True)
    with T.Context(0) as ctx:
        ctx.rdf_init(T.rdf_params(3, 3, 0.03, 4.2), [L] * 3)
        arr = pos if layout == T.TRJ_AOS else np.ascontiguousarray(pos.T)
        ctx.frame(arr, it, [L] * 3, what=T.TRJ_WANT_RDF, layout=layout)
        ctx.frame(arr.astype(np.float32), it, [L] * 3, what=T.TRJ_WANT_RDF, layout=layout)
        g = ctx.rdf_counts()
    o = O.rdf_counts(pos, it, 3, 3, [L] * 3, (1, 1, 1), 4.2, 0.03, method="brute")
    o += O.rdf_counts(pos.astype(np.float32), it, 3, 3, [L] * 3, (1, 1, 1), 4.2, 0.03, method="brute")
    np.testing.assert_array_equal(g, o)


def test_slab_nonperiodic_z():
    rng = np.random.default_rng(5)
    L = 16.0
    pos = rng.random((3000, 3)) * L
    it = rng.integers(1, 3, 3000).astype(np.int32)
    print("case: slab", flush=True)
    g, _ = _gpu_counts([(pos, it, [L] * 3)], 3, 2, 0.02, 5.0, periodic=(1, 1, 0))
    o = _oracle_counts([(pos, it, [L] * 3)], 3, 2, 0.02, 5.0, periodic=(1, 1, 0))
    np.testing.assert_array_equal(g, o)


def test_2d():
    rng = np.random.default_rng(6)
    L = 40.0
    pos = rng.random((3000, 2)) * L
    it = rng.integers(1, 3, 3000).astype(np.int32)
    pos3 = np.concatenate([pos, np.zeros((3000, 1))], 1)
    print("case: 2d", flush=True)
    g, _ = _gpu_counts([(pos, it, [L, L, 1.0])], 2, 2, 0.05, 6.0)
    o = _oracle_counts([(pos3, it, [L, L, 1.0])], 2, 2, 0.05, 6.0)
    np.testing.assert_array_equal(g, o)


def test_small_box_allpairs_path():
    rng = np.random.default_rng(7)
    L = 9.0  # 9/4 < 3 cells: all-pairs kernel
    pos = rng.random((700, 3)) * L
    it = rng.integers(1, 3, 700).astype(np.int32)
    print("case: allpairs", flush=True)
    g, _ = _gpu_counts([(pos, it, [L] * 3)], 3, 2, 0.02, 4.0)
    o = _oracle_counts([(pos, it, [L] * 3)], 3, 2, 0.02, 4.0)
    np.testing.assert_array_equal(g, o)


def test_lattice_262k_bit_exact():
    f = next(W.lattice_262k(nframes=1))
    print("case: lattice_262k", flush=True)
    g, _ = _gpu_counts([(f.pos, f.itype, f.box)], 3, 3, 0.01, 5.0)
    o = _oracle_counts([(f.pos, f.itype, f.box)], 3, 3, 0.01, 5.0, method="cells")
    np.testing.assert_array_equal(g, o)
    assert g.sum() > 5e7


def test_gmix_and_first_minimum_match_oracle():
    fr = [(f.pos, f.itype, f.box) for f in W.lj_binary_4k(nframes=2)]
    print("case: gmix", flush=True)
    with T.Context(0) as ctx:
        st = ctx.rdf_init(T.rdf_params(3, 2, 0.02, 4.0), fr[0][2])
        for pos, it, box in fr:
            ctx.frame(pos, it, box, what=T.TRJ_WANT_RDF)
        g = ctx.rdf_gmix()
        c = ctx.rdf_counts()
        rmin = ctx.rdf_first_minimum()
    L = fr[0][2][0]
    ntype = np.bincount(fr[0][1])[1:]
    o = O.rdf_gmix(3, 2, st.lsmax, 0.02, L ** 3, ntype, 2, c)
    np.testing.assert_allclose(g, o, rtol=1e-13, atol=0)
    assert rmin == pytest.approx(O.first_minimum(3, 2, st.lsmax, 0.02, L ** 3, 4000, c), rel=1e-12)


def test_errors_are_reported():
    print("case: errors", flush=True)
    with T.Context(0) as ctx:
        ctx.rdf_init(T.rdf_params(3, 2, 0.02, 4.0, (1, 1, 0)), [10.0] * 3)
        pos = np.zeros((100, 3))
        ctx.frame(pos, np.full(100, 3, np.int32), [10.0] * 3, what=T.TRJ_WANT_RDF)
        with pytest.raises(T.TrjError) as e:
            ctx.sync()
        assert e.value.code == T.TRJ_EINVAL
        pos[5, 2] = 30.0  # non-periodic axis far outside the box
        ctx.frame(pos, np.ones(100, np.int32), [10.0] * 3, what=T.TRJ_WANT_RDF)
        with pytest.raises(T.TrjError) as e:
            ctx.sync()
        assert e.value.code == T.TRJ_ERANGE
        with pytest.raises(T.TrjError):
            ctx.frame(pos, np.ones(100, np.int32), [10.0] * 3, what=T.TRJ_WANT_SQ)  # no sq_init


def test_deterministic_repeat():
    f = next(W.uneven_batch_frames(1)) if hasattr(W, "uneven_batch_frames") else W.uneven_frame(3, 1e4, 3e4)
    print("case: determinism", flush=True)
    res = []
    for _ in range(2):
        with T.Context(0) as ctx:
            ctx.rdf_init(T.rdf_params(3, f.nsp, 0.01, 4.0), f.box)
            ctx.frame(f.pos, f.itype, f.box, what=T.TRJ_WANT_RDF)
            res.append(ctx.rdf_counts())
    np.testing.assert_array_equal(res[0], res[1])


@pytest.mark.parametrize("L,rc,n", [(30.0, 5.0, 20000), (26.0, 5.0, 12000), (52.0, 4.0, 60000)])
def test_cell_shift_path_bit_exact(L, rc, n):
    """Boxes with >= 5 cells per axis take the neighbour-cell image shift instead
    of the per-pair rule; unwrapped input, three species."""
    rng = np.random.default_rng(int(L * 10))
    pos = rng.random((n, 3)) * L * 3.0 - L  # spans three periods
    it = rng.integers(1, 4, n).astype(np.int32)
    print("case: cell shift", L, rc, n, flush=True)
    g, _ = _gpu_counts([(pos, it, [L] * 3)], 3, 3, 0.01, rc)
    o = _oracle_counts([(pos, it, [L] * 3)], 3, 3, 0.01, rc, method="cells")
    np.testing.assert_array_equal(g, o)


@pytest.mark.parametrize("ndim,L,rc,n", [(3, 230.0, 5.0, 30000), (2, 1800.0, 5.0, 40000)])
def test_sparse_grid_atomic_cell_path(ndim, L, rc, n):
    """Grids of more than 81920 cells (45^3, 359^2) overflow the LDS of the
    per-workgroup cell counts: the cells are sorted by the global-atomic count
    and scatter instead; the counts match all the same."""
    rng = np.random.default_rng(ndim * 100 + n % 97)
    pos = rng.random((n, 3)) * L
    if ndim == 2:
        pos[:, 2] = 0.0
    it = rng.integers(1, 3, n).astype(np.int32)
    box = [L, L, L if ndim == 3 else 1.0]
    print("case: sparse grid", ndim, L, n, flush=True)
    g, _ = _gpu_counts([(pos[:, :ndim] if ndim == 2 else pos, it, box)], ndim, 2, 0.02, rc)
    o = _oracle_counts([(pos, it, box)], ndim, 2, 0.02, rc, method="cells")
    np.testing.assert_array_equal(g, o)
    assert g.sum() > 1000


def test_lds_atomic_probe():
    print("case: LDS-atomic rate probe", flush=True)
    r = T.lds_atomic_rate(0)
    print("LDS atomics/s", r, flush=True)
    assert 1e11 < r < 1e15
