"""GPU parity of F(Q,t) / Fs(Q,t) over time origins against the oracle
(same q-set, same origin bookkeeping; FP64 phases on the device, long double
in the oracle), through the C-ABI."""
import numpy as np
import pytest

from oracle import oracle as O
from trj_analysis_amd import trjhip as T

pytestmark = pytest.mark.gpu


def _traj(nf, n, L, seed, nsp=2):
    rng = np.random.default_rng(seed)
    x = rng.random((n, 3)) * L
    it = rng.integers(1, nsp + 1, n).astype(np.int32)
    frames = []
    for _ in range(nf):
        x = (x + 0.15 * rng.standard_normal((n, 3))) % L
        frames.append(x.copy())
    return frames, it


@pytest.mark.parametrize("nbuffer,jump,ndyn,itlimit", [(3, 2, 5, 5), (1, 1, 8, 8), (4, 1, 4, 3)])
def test_fqt_against_oracle(nbuffer, jump, ndyn, itlimit):
    L = 14.0
    frames, it = _traj(12, 1500, L, seed=nbuffer * 10 + jump)
    qw = [1.1, 2.3, 3.0]
    bsc = [1.0, 0.6]
    print("case: fqt", nbuffer, jump, ndyn, itlimit, flush=True)
    with T.Context(0) as ctx:
        nv, nw = ctx.fqt_init(qw, nbuffer, jump, ndyn, itlimit, [L] * 3, ndim=3, nsp=2, bsc=bsc)
        for x in frames:
            ctx.frame(x, it, [L] * 3, what=T.TRJ_WANT_FQT)
        f, fs, co = ctx.fqt_results()
    kv, bn = O.fqt_vectors(3, [L] * 3, qw)
    assert nv == len(kv) and list(nw) == [int(np.sum(bn == l)) for l in range(3)]
    tot, tots, oco = O.fqt(frames, it, 3, 2, bsc, [L] * 3, qw, nbuffer, jump, ndyn, itlimit)
    assert co == oco > 0
    den = np.array([1500.0 * co * w for w in nw])[:, None]
    np.testing.assert_allclose(f, tot / den, rtol=1e-9, atol=1e-11)
    np.testing.assert_allclose(fs, tots / den, rtol=1e-9, atol=1e-11)
    assert np.all(np.abs(fs[:, 0] - 1.0) < 1e-12)          # lag 0: Fs = 1


def test_fqt_with_rdf_in_one_frame_call():
    L = 12.0
    frames, it = _traj(6, 900, L, seed=5)
    print("case: fqt with rdf", flush=True)
    with T.Context(0) as ctx:
        ctx.rdf_init(T.rdf_params(3, 2, 0.05, 4.0), [L] * 3)
        ctx.fqt_init([1.5], 2, 1, 3, 3, [L] * 3, ndim=3, nsp=2)
        for x in frames:
            ctx.frame(x, it, [L] * 3, what=T.TRJ_WANT_RDF | T.TRJ_WANT_FQT)
        f, fs, co = ctx.fqt_results()
        counts = ctx.rdf_counts()
    tot, tots, oco = O.fqt(frames, it, 3, 2, [1.0, 1.0], [L] * 3, [1.5], 2, 1, 3, 3)
    kv, _ = O.fqt_vectors(3, [L] * 3, [1.5])
    assert co == oco
    np.testing.assert_allclose(fs, tots / (900.0 * co * len(kv)), rtol=1e-9)
    assert counts.sum() > 0


def test_fqt_only_2d_host_frame():
    """An F(Q,t)-only frame on a 2D context reads r(2,N) from the host: the
    staging copy is sized by the F(Q,t) module's ndim (regression: it used to
    take 3 coordinates per atom from a 2-coordinate array)."""
    L = 20.0
    rng = np.random.default_rng(21)
    n = 1200
    x = rng.random((n, 2)) * L
    it = rng.integers(1, 3, n).astype(np.int32)
    frames = []
    for _ in range(7):
        x = (x + 0.2 * rng.standard_normal((n, 2))) % L
        frames.append(x.copy())
    qw = [0.9, 1.7]
    print("case: fqt-only 2D host frame", flush=True)
    with T.Context(0) as ctx:
        nv, nw = ctx.fqt_init(qw, 2, 2, 4, 4, [L, L, 1.0], ndim=2, nsp=2, bsc=[1.0, 0.5])
        for f in frames:
            ctx.frame(f, it, [L, L, 1.0], what=T.TRJ_WANT_FQT)
        f2, fs2, co = ctx.fqt_results()
    padded = [np.c_[f, np.zeros(n)] for f in frames]       # the oracle reads 3 values per atom
    tot, tots, oco = O.fqt(padded, it, 2, 2, [1.0, 0.5], [L, L, 1.0], qw, 2, 2, 4, 4)
    assert co == oco > 0
    den = np.array([n * co * w for w in nw], dtype=np.float64)[:, None]
    np.testing.assert_allclose(f2, tot / den, rtol=1e-9, atol=1e-11)
    np.testing.assert_allclose(fs2, tots / den, rtol=1e-9, atol=1e-11)


def test_fqt_ndim_mismatch_rejected():
    """S(Q) set up in 2D and F(Q,t) in 3D: a frame asking for both is refused
    with TRJ_EINVAL before any input is read."""
    L = 15.0
    rng = np.random.default_rng(3)
    x = rng.random((500, 2)) * L
    it = np.ones(500, dtype=np.int32)
    print("case: fqt ndim mismatch", flush=True)
    with T.Context(0) as ctx:
        ctx.sq_init(T.sq_params(2, 1, 3.0), [L, L, 1.0])
        ctx.fqt_init([1.0], 1, 1, 2, 2, [L] * 3, ndim=3, nsp=1)
        with pytest.raises(T.TrjError) as e:
            ctx.frame(x, it, [L, L, 1.0], what=T.TRJ_WANT_SQ | T.TRJ_WANT_FQT)
        assert e.value.code == T.TRJ_EINVAL
        assert "different ndim" in str(e.value)
