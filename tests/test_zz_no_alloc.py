"""trj_frame_device_noalloc (RDF of a device-resident frame with no
allocation: per-frame buffers in the library's static scratch) against
trj_frame_device, through the C-ABI: counts bit-exact after each of two calls
in a row, the second with a smaller input (and another dtype and layout);
the all-pairs path; refusals checked on the host before any launch."""
import numpy as np
import pytest

from oracle import oracle as O
from trj_analysis_amd import trjhip as T

pytestmark = pytest.mark.gpu


def _frame(n, L, nsp, seed):
    rng = np.random.default_rng(seed)
    pos = rng.random((n, 3)) * L
    it = rng.integers(1, nsp + 1, n).astype(np.int32)
    return pos, it


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def test_noalloc_matches_frame_device_two_calls():
    import torch
    nsp, rc, dr = 3, 5.0, 0.01
    L1 = (300_000 / 0.85) ** (1 / 3)
    p1, t1 = _frame(300_000, L1, nsp, 1)
    L2 = (40_000 / 0.85) ** (1 / 3)
    p2, t2 = _frame(40_000, L2, nsp, 2)
    # second frame: float32, structure-of-arrays (x[N] y[N] z[N])
    p2_soa = np.ascontiguousarray(p2.T.astype(np.float32))
    d1, dt1 = _dev(p1), _dev(t1)
    d2, dt2 = _dev(p2_soa), _dev(t2)
    torch.cuda.synchronize()
    calls = [(d1, T.TRJ_F64, T.TRJ_AOS, dt1, len(t1), [L1] * 3), (d2, T.TRJ_F32, T.TRJ_SOA, dt2, len(t2), [L2] * 3)]
    print("case: noalloc two calls", [c[4] for c in calls], flush=True)
    with T.Context(0) as a, T.Context(0) as b:
        a.rdf_init(T.rdf_params(3, nsp, dr, rc), [L1] * 3)
        b.rdf_init(T.rdf_params(3, nsp, dr, rc), [L1] * 3)
        for k, (pp, dty, lay, tt, n, box) in enumerate(calls):
            a.frame_device(pp.data_ptr(), dty, lay, tt.data_ptr(), n, box, what=T.TRJ_WANT_RDF)
            b.frame_device_noalloc(pp.data_ptr(), dty, lay, tt.data_ptr(), n, box)
            ca, cb = a.rdf_counts(), b.rdf_counts()
            print("call", k, "pairs", int(cb.sum()), flush=True)
            np.testing.assert_array_equal(cb, ca)
            assert cb.sum() > 0
        np.testing.assert_allclose(b.rdf_gmix(2), a.rdf_gmix(2), rtol=0, atol=0)
    # the smaller frame alone against the oracle
    with T.Context(0) as b:
        b.rdf_init(T.rdf_params(3, nsp, dr, rc), [L2] * 3)
        b.frame_device_noalloc(d2.data_ptr(), T.TRJ_F32, T.TRJ_SOA, dt2.data_ptr(), len(t2), [L2] * 3)
        got = b.rdf_counts()
    want = O.rdf_counts(p2.astype(np.float32), t2, 3, nsp, [L2] * 3, (1, 1, 1), rc, dr, method="cells")
    np.testing.assert_array_equal(got, want)


def test_noalloc_allpairs_path_and_slab():
    """A box with fewer than 3 cells per axis (the all-pairs kernel), then a
    z-confined slab (the cell path without an image on z)."""
    import torch
    p, t = _frame(3000, 9.0, 2, 5)
    d, dt = _dev(p), _dev(t)
    ps, ts = _frame(20000, 30.0, 2, 6)
    ds, dts = _dev(ps), _dev(ts)
    torch.cuda.synchronize()
    print("case: noalloc all-pairs and slab", flush=True)
    for (pp, tt, n, box, per) in ((d, dt, 3000, [9.0] * 3, (True, True, True)),
                                  (ds, dts, 20000, [30.0] * 3, (True, True, False))):
        with T.Context(0) as a, T.Context(0) as b:
            for c in (a, b):
                c.rdf_init(T.rdf_params(3, 2, 0.02, 4.0, per), box)
            a.frame_device(pp.data_ptr(), T.TRJ_F64, T.TRJ_AOS, tt.data_ptr(), n, box, what=T.TRJ_WANT_RDF)
            b.frame_device_noalloc(pp.data_ptr(), T.TRJ_F64, T.TRJ_AOS, tt.data_ptr(), n, box)
            np.testing.assert_array_equal(b.rdf_counts(), a.rdf_counts())


def test_noalloc_refusals():
    """Refused on the host before anything is launched: a frame whose buffers
    exceed the scratch (TRJ_ELIMIT), S(Q) asked for, and a charged context
    (TRJ_EINVAL)."""
    import torch
    n = 2_000_000
    L = (n / 0.85) ** (1 / 3)
    p, t = _frame(n, L, 2, 7)
    d, dt = _dev(p), _dev(t)
    torch.cuda.synchronize()
    print("case: noalloc refusals", flush=True)
    with T.Context(0) as b:
        b.rdf_init(T.rdf_params(3, 2, 0.01, 4.0), [L] * 3)
        with pytest.raises(T.TrjError) as e:
            b.frame_device_noalloc(d.data_ptr(), T.TRJ_F64, T.TRJ_AOS, dt.data_ptr(), n, [L] * 3)
        assert e.value.code == T.TRJ_ELIMIT and "scratch" in str(e.value)
        assert not b.rdf_counts().any()
        rc = b._L.trj_frame_device_noalloc(b._h, T.C.c_void_p(d.data_ptr()), T.TRJ_F64, T.TRJ_AOS,
                                           T.C.c_void_p(dt.data_ptr()), n, T._ptr(T._box3([L] * 3), T.C.c_double),
                                           T.TRJ_WANT_RDF | T.TRJ_WANT_SQ)
        assert rc == T.TRJ_EINVAL
    with T.Context(0) as b:
        b.rdf_init(T.rdf_params(3, 2, 0.01, 4.0, flags=T.TRJ_RDF_CHARGE), [L] * 3)
        with pytest.raises(T.TrjError) as e:
            b.frame_device_noalloc(d.data_ptr(), T.TRJ_F64, T.TRJ_AOS, dt.data_ptr(), 1000, [L] * 3)
        assert e.value.code == T.TRJ_EINVAL
