"""Batched S(Q) (trj_batch_sq_run): every frame's shell sums against the
per-frame path on the same box (trj_sq_init on the half-space grid, one
trj_frame, trj_sq_shells) and, for a small frame, against the oracle's
long-double phase sums; with and without graph replay, over several groups."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from trj_analysis_amd import trjhip as T

pytestmark = pytest.mark.gpu


def _frames(seed, nf, ndim=3, nmin=2e3, nmax=6e4):
    """Uneven frames: non-cubic boxes, unwrapped float32 positions, 1-4
    species, per-frame scattering lengths."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nf):
        n = int(10 ** rng.uniform(np.log10(nmin), np.log10(nmax)))
        nsp = int(rng.integers(1, 5))
        box = rng.uniform(12.0, 30.0, 3)
        if ndim == 2:
            box[2] = 1.0
        pos = (rng.random((n, ndim)) * 1.5 - 0.25) * box[:ndim]
        it = rng.integers(1, nsp + 1, n).astype(np.int32)
        bsc = rng.uniform(-0.5, 1.5, nsp)
        out.append((pos.astype(np.float32), it, box, nsp, bsc))
    return out


def _batch(frames, ndim, grid, prec, graph=True, fpl=0, qmax=1e9):
    pos = torch.from_numpy(np.concatenate([f[0] for f in frames])).cuda()
    it = torch.from_numpy(np.concatenate([f[1] for f in frames])).cuda()
    fa = T.batch_frames((len(f[1]), f[3], f[2], (1, 1, 1)) for f in frames)
    b = T.Batch(T.batch_params(ndim, 0.01, 4.0, frames_per_launch=fpl, use_graph=graph))
    q = T.batch_sq_params(grid, prec, qmax)
    out, ms = b.sq_run(q, fa, pos.data_ptr(), T.TRJ_F32, it.data_ptr(), bsc=[f[4] for f in frames])
    out2, _ = b.sq_run(q, fa, pos.data_ptr(), T.TRJ_F32, it.data_ptr(), bsc=[f[4] for f in frames])
    b.close()
    assert ms > 0
    return out, out2


def _per_frame(frames, ndim, grid, prec, qmax=1e9):
    res = []
    with T.Context(0) as ctx:
        for pos, it, box, nsp, bsc in frames:
            ctx.sq_init(T.sq_params(ndim, nsp, qmax, bsc=list(bsc), qset=T.TRJ_QSET_HALFSPACE, grid=grid,
                                    precision=prec), box)
            ctx.frame(pos, it, box, what=T.TRJ_WANT_SQ)
            res.append(ctx.sq_shells())
    return res


def _compare(got, want, rtol):
    for k, ((sqf, sqfp, nq), (wf, wfp, wnq)) in enumerate(zip(got, want)):
        np.testing.assert_array_equal(nq, wnq)
        assert sqf.shape == wf.shape and sqfp.shape == wfp.shape
        m = wnq > 0
        np.testing.assert_allclose(sqf[m], wf[m], rtol=rtol, err_msg=f"frame {k} sqf")
        np.testing.assert_allclose(sqfp[:, m], wfp[:, m], rtol=rtol, err_msg=f"frame {k} sqfp")
        assert np.all(sqf[~m] == 0.0)


@pytest.mark.parametrize("graph,fpl", [(True, 3), (False, 0), (True, 64)])
def test_batch_sq_fp64_matches_per_frame(graph, fpl):
    frames = _frames(11, 7)
    print("case: batch sq fp64 graph", graph, "fpl", fpl, flush=True)
    got, got2 = _batch(frames, 3, (32, 32, 32), T.TRJ_PREC_FP64, graph, fpl)
    want = _per_frame(frames, 3, (32, 32, 32), "fp64")
    _compare(got, want, 1e-10)
    _compare(got2, want, 1e-10)


def test_batch_sq_qmax_cut_and_2d():
    frames = _frames(12, 5, ndim=2)
    print("case: batch sq 2d, qmax cut", flush=True)
    got, _ = _batch(frames, 2, (32, 32, 2), T.TRJ_PREC_FP64, fpl=2, qmax=3.0)
    want = _per_frame(frames, 2, (32, 32, 2), "fp64", qmax=3.0)
    _compare(got, want, 1e-10)


@pytest.mark.parametrize("prec,tol", [(T.TRJ_PREC_FP16X3, 1e-5), (T.TRJ_PREC_FP32, 1e-5)])
def test_batch_sq_reduced_precision(prec, tol):
    """Split-16-bit and FP32 contractions of the batch against the per-frame
    FP64 path, per shell."""
    frames = _frames(13, 6, nmin=5e3, nmax=4e4)
    print("case: batch sq precision", prec, flush=True)
    got, _ = _batch(frames, 3, (32, 32, 32), prec, fpl=4)
    want = _per_frame(frames, 3, (32, 32, 32), "fp64")
    worst = 0.0
    for (sqf, sqfp, nq), (wf, wfp, wnq) in zip(got, want):
        np.testing.assert_array_equal(nq, wnq)
        m = wf > 0
        worst = max(worst, float((np.abs(sqf[m] - wf[m]) / wf[m]).max()))
    print("worst shell rel", worst, flush=True)
    assert worst < tol


@pytest.mark.parametrize("prec", [T.TRJ_PREC_FP64, T.TRJ_PREC_FP16X3])
def test_batch_sq_anchor_table_on_off(monkeypatch, prec):
    """The batch's contraction with its table of phase anchors (the default)
    and without (TRJ_SQ_ANCHOR_TABLE=0): the table holds the values the
    kernel would evaluate, so the partials are equal and the shells agree to
    the order of their FP64 atomics; groups of 3 frames over a shared
    buffer (the table covers each group's entries only)."""
    frames = _frames(15, 7, nmin=3e3, nmax=5e4)
    res = {}
    for table in ("1", "0"):
        monkeypatch.setenv("TRJ_SQ_ANCHOR_TABLE", table)
        print("case: batch sq anchor table", table, "precision", prec, flush=True)
        res[table] = _batch(frames, 3, (32, 32, 32), prec, fpl=3)[0]
    for k, ((sqf, sqfp, nq), (wf, wfp, wnq)) in enumerate(zip(res["1"], res["0"])):
        np.testing.assert_array_equal(nq, wnq)
        np.testing.assert_allclose(sqf, wf, rtol=1e-13, atol=0, err_msg=f"frame {k} sqf")
        np.testing.assert_allclose(sqfp, wfp, rtol=1e-13, atol=0, err_msg=f"frame {k} sqfp")


def test_batch_sq_against_oracle():
    """A small frame inside a batch: shells summed from the oracle's long-double
    rho of every half-space vector (the origin left out), shells by the plan's
    FP64 rule, vector counts equal to the per-frame plan's nq."""
    frames = _frames(14, 3, nmin=1500, nmax=3000)
    grid = (8, 8, 8)
    print("case: batch sq oracle", flush=True)
    got, _ = _batch(frames, 3, grid, T.TRJ_PREC_FP64, fpl=2)
    for k, (pos, it, box, nsp, bsc) in enumerate(frames):
        wrapped = np.mod(pos.astype(np.float64), box)
        kv = np.array([(a, b, c) for a in range(-4, 4) for b in range(-4, 4) for c in range(0, 4)
                       if (a, b, c) != (0, 0, 0)], dtype=np.int32)
        rho = O.sq_rho(wrapped, it, nsp, box, kv)
        fk = 2.0 * np.pi / box
        dq = fk.min()
        q = np.sqrt((kv[:, 0] * fk[0]) ** 2 + (kv[:, 1] * fk[1]) ** 2 + (kv[:, 2] * fk[2]) ** 2)
        sh = np.floor(q / dq + 0.5).astype(int)
        sqf, sqfp, nq = got[k]
        info, pnq = T.sq_plan(T.sq_params(3, nsp, 1e9, qset=T.TRJ_QSET_HALFSPACE, grid=grid), box)
        np.testing.assert_array_equal(nq, pnq)
        wf = np.zeros_like(sqf)
        wfp = np.zeros_like(sqfp)
        tot = rho @ np.asarray(bsc)
        keep = (sh >= 1) & (sh <= info.nqmax)
        np.add.at(wf, sh[keep], np.abs(tot[keep]) ** 2)
        for s in range(nsp):
            np.add.at(wfp[s], sh[keep], np.abs(rho[keep, s]) ** 2)
        m = nq > 0
        np.testing.assert_allclose(sqf[m], wf[m], rtol=1e-10)
        np.testing.assert_allclose(sqfp[:, m], wfp[:, m], rtol=1e-10)


def test_batch_sq_rejects_bad_types():
    frames = _frames(15, 2)
    pos, it, box, nsp, bsc = frames[1]
    bad = it.copy()
    bad[3] = nsp + 1
    frames[1] = (pos, bad, box, nsp, bsc)
    print("case: batch sq bad types", flush=True)
    with pytest.raises(T.TrjError):
        _batch(frames, 3, (16, 16, 16), T.TRJ_PREC_FP64)
