"""Pins of the oracle's F(Q,t): the q-set against an FP64 enumeration,
identical frames, a uniform drift (closed forms) and origin counts worked by
hand from rtcorr's rules."""
import numpy as np

from oracle import oracle as O


def test_vector_set_shells():
    box = [12.0, 13.0, 14.0]
    qw = [1.5, 2.6]
    kv, bn = O.fqt_vectors(3, box, qw)
    fk = 2 * np.pi / np.array(box)
    dq = fk.min()
    q = np.sqrt(((kv * fk) ** 2).sum(1))
    for l, t in enumerate(qw):
        m = bn == l
        assert np.all(np.abs(q[m] - t) < dq / 2 + 1e-5)
    # every vector well inside a window is present (FP64 enumeration)
    r = np.arange(0, 8)
    g = np.array(np.meshgrid(r, r, r, indexing="ij")).reshape(3, -1).T
    g = g[g.sum(1) > 0]
    qq = np.sqrt(((g * fk) ** 2).sum(1))
    have = {tuple(v) for v in kv}
    for t in qw:
        inside = g[np.abs(qq - t) < dq / 2 - 1e-4]
        assert all(tuple(v) in have for v in inside)
    assert len(kv) == len(have)


def _frames(nf, n, L, seed, drift=None):
    rng = np.random.default_rng(seed)
    x0 = rng.random((n, 3)) * L
    out = []
    for f in range(nf):
        out.append(x0 + (0.0 if drift is None else f * np.asarray(drift)))
    return out


def test_identical_frames():
    L = 10.0
    fr = _frames(7, 60, L, 1)
    it = np.ones(60, dtype=np.int32)
    tot, tots, co = O.fqt(fr, it, 3, 1, [1.0], [L] * 3, [1.3], 2, 1, 3, 3)
    kv, bn = O.fqt_vectors(3, [L] * 3, [1.3])
    assert co == 2                       # buffer 1: frames 1-4, buffer 2: 2-5 complete
    np.testing.assert_array_equal(tots[0], co * len(kv) * 60.0)
    assert np.allclose(tot[0], tot[0][0]) and tot[0][0] > 0


def test_uniform_drift_closed_form():
    L = 10.0
    d = np.array([0.13, -0.07, 0.21])
    fr = _frames(6, 40, L, 2, drift=d)
    it = np.ones(40, dtype=np.int32)
    qw = [1.9]
    tot, tots, co = O.fqt(fr, it, 3, 1, [1.0], [L] * 3, qw, 1, 1, 4, 4)
    kv, _ = O.fqt_vectors(3, [L] * 3, qw)
    assert co == 1
    for lag in range(4):
        ph = 2 * np.pi * (kv @ d) * lag / L
        np.testing.assert_allclose(tots[0, lag], 40.0 * np.cos(ph).sum(), rtol=1e-12, atol=1e-9)
        x0 = fr[0]
        rho0 = np.exp(2j * np.pi * (x0 @ (kv.T / L))).sum(0)
        np.testing.assert_allclose(tot[0, lag], (np.abs(rho0) ** 2 * np.cos(ph)).sum(), rtol=1e-10, atol=1e-7)


def test_origin_counts_by_hand():
    # nbuffer 2, jump 1, itlimit 3: buffer 1 completes at frames 4 and 8,
    # buffer 2 (origin 2) at frame 5 -> 3 completed origins in 8 frames
    fr = _frames(8, 5, 8.0, 3)
    _, _, co = O.fqt(fr, np.ones(5, dtype=np.int32), 3, 1, [1.0], [8.0] * 3, [2.0], 2, 1, 3, 3)
    assert co == 3
    _, _, co = O.fqt(fr, np.ones(5, dtype=np.int32), 3, 1, [1.0], [8.0] * 3, [2.0], 3, 2, 2, 2)
    # origins 1, 3, 5, restart 2 frames after the origin: buffer 1 completes at
    # frames 3 and 6, buffer 2 at 5 and 8, buffer 3 at 7: 5 completed
    assert co == 5
