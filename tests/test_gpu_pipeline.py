"""GPU: the overlapped frame pipeline (RDF on its own stream beside the S(Q)
contraction) gives the same results as the serial one, over several frames in
a row and with frames that ask for only one of the two: RDF counts bit for
bit; S(Q) shells to 1e-13 relative (the shell sums use FP64 atomics, whose
order changes the last bits from run to run in either mode)."""
import numpy as np
import pytest

from trj_analysis_amd import trjhip as T
from trj_analysis_amd import workloads as W

pytestmark = pytest.mark.gpu


def _run(mode, frames, precision, whats):
    wl = W.LJ_BINARY_4K
    box = frames[0].box
    with T.Context(0) as ctx:
        ctx.set_pipeline(mode)
        ctx.rdf_init(T.rdf_params(3, 2, wl.deltar, wl.rcrdf), box)
        ctx.sq_init(T.sq_params(3, 2, wl.qmax, wl.qmin, wl.bsc, precision=precision), box)
        for f, what in zip(frames, whats):
            ctx.frame(f.pos, f.itype, f.box, what=what)
        counts = ctx.rdf_counts()
        sqf, sqfp, nq = ctx.sq_shells()
        ms = ctx.last_timings()
    return counts, sqf, sqfp, ms


@pytest.mark.parametrize("precision", ["fp64", "bf16x3"])
def test_overlap_matches_serial(precision):
    frames = list(W.lj_binary_4k(nframes=4))
    both = T.TRJ_WANT_RDF | T.TRJ_WANT_SQ
    whats = [both, both, T.TRJ_WANT_RDF, both]
    print(f"case: lj_binary_4k x4 {precision} serial vs overlap", flush=True)
    c0, s0, p0, _ = _run(T.TRJ_PIPE_SERIAL, frames, precision, whats)
    c1, s1, p1, ms = _run(T.TRJ_PIPE_OVERLAP, frames, precision, whats)
    np.testing.assert_array_equal(c0, c1)
    np.testing.assert_allclose(s1, s0, rtol=1e-13, atol=0)
    np.testing.assert_allclose(p1, p0, rtol=1e-13, atol=0)
    assert c0.sum() > 1_000_000 and ms[0] > 0.0 and ms[1] > 0.0


def test_overlap_large_frame_matches_serial():
    f = next(W.lattice_262k(nframes=1, m=40))
    wl = W.LATTICE_262K
    out = []
    print("case: lattice 64000 atoms serial vs overlap", flush=True)
    for mode in (T.TRJ_PIPE_SERIAL, T.TRJ_PIPE_OVERLAP):
        with T.Context(0) as ctx:
            ctx.set_pipeline(mode)
            ctx.rdf_init(T.rdf_params(3, 3, wl.deltar, wl.rcrdf), f.box)
            ctx.sq_init(T.sq_params(3, 3, 8.0, None, wl.bsc), f.box)
            for _ in range(3):
                ctx.frame(f.pos, f.itype, f.box)
            out.append((ctx.rdf_counts(), ctx.sq_shells()[0]))
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_allclose(out[1][1], out[0][1], rtol=1e-13, atol=0)


def test_bad_pipeline_mode_rejected():
    with T.Context(0) as ctx:
        with pytest.raises(T.TrjError):
            ctx.set_pipeline(7)
