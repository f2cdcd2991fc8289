"""GPU: the accumulated state of a context saved after some frames and loaded
into another context that processes the rest equals one context that
processed every frame (RDF counts bit for bit, S(Q) shell sums to 1e-13
relative: the shell sums use FP64 atomics); adding a blob to a context's own
sums merges runs; blobs of other plans are refused."""
import numpy as np
import pytest

from trj_analysis_amd import trjhip as T
from trj_analysis_amd import workloads as W

pytestmark = pytest.mark.gpu

WL = W.LJ_BINARY_4K


def _ctx(box, qmax=WL.qmax):
    ctx = T.Context(0)
    ctx.rdf_init(T.rdf_params(3, 2, WL.deltar, WL.rcrdf), box)
    ctx.sq_init(T.sq_params(3, 2, qmax, None, WL.bsc), box)
    return ctx


def _results(ctx):
    sqf, sqfp, _ = ctx.sq_shells()
    cols, mix = ctx.sq_columns()
    return ctx.rdf_counts(), sqf, sqfp, ctx.rdf_gmix(), cols


def _same(a, b):
    np.testing.assert_array_equal(a[0], b[0])
    for x, y in zip(a[1:], b[1:]):
        np.testing.assert_allclose(x, y, rtol=1e-13, atol=0)


def test_save_resume_and_merge_equal_one_run():
    frames = list(W.lj_binary_4k(nframes=4))
    box = frames[0].box
    print("case: lj_binary_4k x4, checkpoint after 2", flush=True)
    with _ctx(box) as full:
        for f in frames:
            full.frame(f.pos, f.itype, f.box)
        want = _results(full)
    with _ctx(box) as first:
        for f in frames[:2]:
            first.frame(f.pos, f.itype, f.box)
        blob = first.state_save()
    # resume: load, then the remaining frames
    with _ctx(box) as resumed:
        resumed.state_load(blob, T.TRJ_STATE_REPLACE)
        for f in frames[2:]:
            resumed.frame(f.pos, f.itype, f.box)
        _same(_results(resumed), want)
    # merge: the remaining frames first, then add the saved part
    with _ctx(box) as merged:
        for f in frames[2:]:
            merged.frame(f.pos, f.itype, f.box)
        merged.state_load(blob, T.TRJ_STATE_ADD)
        _same(_results(merged), want)
    # a loaded state alone gives the first run's normalised output
    with _ctx(box) as loaded:
        loaded.state_load(blob)
        with _ctx(box) as again:
            for f in frames[:2]:
                again.frame(f.pos, f.itype, f.box)
            _same(_results(loaded), _results(again))


def test_foreign_state_refused():
    f = next(W.lj_binary_4k(nframes=1))
    with _ctx(f.box) as a:
        a.frame(f.pos, f.itype, f.box)
        blob = a.state_save()
    with _ctx(f.box, qmax=WL.qmax / 2) as b:
        with pytest.raises(T.TrjError):
            b.state_load(blob)
        with pytest.raises(T.TrjError):
            b.state_load(b"not a state blob at all........" * 8)
