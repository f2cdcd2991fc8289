"""GPU parity of the number-fluctuation sums (s2D3_sh / s2D2_sh): the int64
sums of N_k and N_k^2 around the library's fixed points, bit-exact against the
oracle run on the same points, through the C-ABI."""
import numpy as np
import pytest

from oracle import oracle as O
from trj_analysis_amd import trjhip as T
from trj_analysis_amd import workloads as W

pytestmark = pytest.mark.gpu


def _run(frames, ndim, nsp, box, periodic, rc, dr, nrandom, seed=12345):
    with T.Context(0) as ctx:
        st = ctx.rdf_init(T.rdf_params(ndim, nsp, dr, rc, periodic), box)
        ctx.rdf_sampling(nrandom, seed)
        pts = ctx.rdf_sampling_points()
        for pos, it in frames:
            ctx.frame(pos, it, box, what=T.TRJ_WANT_RDF)
        gi, g2i = ctx.rdf_s2n_sums()
        s2 = ctx.rdf_s2n()
        return st.lsmax, pts, gi, g2i, s2


def _oracle(frames, pts, ndim, box, periodic, rc, dr, ls):
    gi = g2i = None
    for pos, _ in frames:
        p3 = pos if pos.shape[1] == 3 else np.concatenate([pos, np.zeros((len(pos), 1))], 1)
        gi, g2i = O.s2n_sums(pts, p3, ndim, box, periodic, rc, dr, lsmax=ls, gi=gi, g2i=g2i)
    return gi, g2i


def test_cells_lj_binary():
    fr = [(f.pos, f.itype) for f in W.lj_binary_4k(nframes=2)]
    box = next(W.lj_binary_4k(nframes=1)).box
    print("case: s2n cells lj_binary_4k", flush=True)
    ls, pts, gi, g2i, s2 = _run(fr, 3, 2, box, (1, 1, 1), 4.0, 0.02, 300)
    assert pts.shape == (300, 3) and pts.min() > 0.0 and pts.max() <= 1.0
    ogi, og2i = _oracle(fr, pts, 3, box, (1, 1, 1), 4.0, 0.02, ls)
    np.testing.assert_array_equal(gi, ogi)
    np.testing.assert_array_equal(g2i, og2i)
    m = 2.0 * 300
    i = np.arange(1, ls - 1)
    np.testing.assert_allclose(s2, g2i[i] / m - (gi[i] / m) ** 2, rtol=1e-15, atol=1e-12)
    assert gi[-1] > 0


@pytest.mark.parametrize("case", ["allpairs", "slab", "2d"])
def test_geometries(case):
    rng = np.random.default_rng({"allpairs": 1, "slab": 2, "2d": 3}[case])
    if case == "allpairs":
        L, n, rc, dr, ndim, per = 7.0, 400, 3.4, 0.05, 3, (1, 1, 1)
    elif case == "slab":
        L, n, rc, dr, ndim, per = 16.0, 3000, 4.0, 0.05, 3, (1, 1, 0)
    else:
        L, n, rc, dr, ndim, per = 40.0, 3000, 6.0, 0.05, 2, (1, 1, 1)
    box = [L, L, L if ndim == 3 else 1.0]
    it = rng.integers(1, 3, n).astype(np.int32)
    frames = []
    for _ in range(2):
        pos = rng.random((n, 3)) * L
        frames.append((pos[:, :ndim] if ndim == 2 else pos, it))
    print("case: s2n", case, flush=True)
    ls, pts, gi, g2i, _ = _run(frames, ndim, 2, box, per, rc, dr, 100, seed=99)
    ogi, og2i = _oracle(frames, pts, ndim, box, per, rc, dr, ls)
    np.testing.assert_array_equal(gi, ogi)
    np.testing.assert_array_equal(g2i, og2i)


def test_sampling_checks():
    print("case: s2n checks", flush=True)
    with T.Context(0) as ctx:
        with pytest.raises(T.TrjError):
            ctx.rdf_sampling(10)                       # before rdf_init
        ctx.rdf_init(T.rdf_params(3, 1, 0.05, 3.0), [10.0] * 3)
        ctx.rdf_sampling(10, 1)
        a = ctx.rdf_sampling_points()
        ctx.rdf_sampling(10, 1)
        np.testing.assert_array_equal(ctx.rdf_sampling_points(), a)   # seeded
        ctx.rdf_sampling(10, 2)
        assert not np.array_equal(ctx.rdf_sampling_points(), a)
        with pytest.raises(T.TrjError):
            ctx.state_save()
