"""GPU parity of the S(Q) path against the long-double oracle."""
import numpy as np
import pytest

from oracle import oracle as O
from trj_analysis_amd import trjhip as T
from trj_analysis_amd import workloads as W

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    m = np.abs(b) > 0
    return float(np.max(np.abs(a[m] - b[m]) / np.abs(b[m]))) if m.any() else 0.0


def _wrapped(f):
    return O.wrap(f.pos, f.box, (1, 1, 1))


def test_lj_binary_shells_fp64():
    frames = list(W.lj_binary_4k(nframes=2))
    wl = W.LJ_BINARY_4K
    print("case: lj_binary shells", flush=True)
    with T.Context(0) as ctx:
        info = ctx.sq_init(T.sq_params(3, 2, wl.qmax, wl.qmin, wl.bsc), frames[0].box)
        for f in frames:
            ctx.frame(f.pos, f.itype, f.box, what=T.TRJ_WANT_SQ)
        sqf, sqfp, nq = ctx.sq_shells()
        cols, mix = ctx.sq_columns()
    g = O.sq_setup(3, frames[0].box, wl.qmax, wl.qmax)
    osqf = np.zeros(g.nqmax + 1)
    osqfp = np.zeros((2, g.nqmax + 1))
    for f in frames:
        O.sq_frame(_wrapped(f), f.itype, 3, 2, f.box, g, wl.bsc, osqf, osqfp)
    assert info.nqmax == g.nqmax == 27 and info.nqmin == g.nqmin == 26
    print("max rel sqf", _rel(sqf, osqf), "sqfp", _rel(sqfp, osqfp), flush=True)
    assert _rel(sqf, osqf) <= 1e-12
    assert _rel(sqfp, osqfp) <= 1e-12
    np.testing.assert_array_equal(nq, O.sq_nq(3, g))
    ntype = np.bincount(frames[0].itype)[1:]
    ocols, omix = O.sq_columns(2, g, 4000, ntype, wl.bsc, 2, osqf, osqfp, nq)
    np.testing.assert_allclose(cols, ocols, rtol=1e-11)
    np.testing.assert_allclose(mix, omix, rtol=1e-11)


def test_rho_per_vector_fp64():
    f = next(W.lj_binary_4k(nframes=1))
    print("case: rho per vector", flush=True)
    with T.Context(0) as ctx:
        ctx.sq_init(T.sq_params(3, 2, 10.0, 10.0, [1.0, 0.5]), f.box)
        ctx.frame(f.pos, f.itype, f.box, what=T.TRJ_WANT_SQ)
        g = O.sq_setup(3, f.box, 10.0, 10.0)
        rng = np.random.default_rng(0)
        kv = []
        while len(kv) < 300:
            v = rng.integers(0, g.kext[0], 3)
            if O.sq_member(3, g, *v) >= 0:
                kv.append(v)
        kv = np.array(kv, dtype=np.int32)
        rho = ctx.sq_rho(kv)
    orho = O.sq_rho(_wrapped(f), f.itype, 2, f.box, kv)
    err = np.abs(rho - orho).max()
    print("max |drho|", err, flush=True)
    assert err < 1e-12


def test_axis_only_mode():
    f = next(W.lj_binary_4k(nframes=1))
    print("case: axis only", flush=True)
    with T.Context(0) as ctx:
        info = ctx.sq_init(T.sq_params(3, 2, 6.0, 0.3, [1.0, 0.5]), f.box)
        ctx.frame(f.pos, f.itype, f.box, what=T.TRJ_WANT_SQ)
        sqf, sqfp, nq = ctx.sq_shells()
    g = O.sq_setup(3, f.box, 6.0, 0.3)
    assert info.nqmin == g.nqmin <= 1
    osqf, osqfp = O.sq_frame(_wrapped(f), f.itype, 3, 2, f.box, g, [1.0, 0.5])
    assert _rel(sqf, osqf) <= 1e-12 and _rel(sqfp, osqfp) <= 1e-12
    assert np.all(nq[1:] == 3)


def test_halfspace_grid_rho():
    rng = np.random.default_rng(8)
    n, L = 20000, 28.0
    pos = rng.random((n, 3)) * L
    it = rng.integers(1, 4, n).astype(np.int32)
    print("case: halfspace", flush=True)
    with T.Context(0) as ctx:
        info = ctx.sq_init(T.sq_params(3, 3, 1e9, bsc=[1.0, 0.6, -0.3], qset=T.TRJ_QSET_HALFSPACE,
                                       grid=(16, 16, 16)), [L] * 3)
        ctx.frame(pos, it, [L] * 3, what=T.TRJ_WANT_SQ)
        kv = np.stack([rng.integers(-8, 8, 200), rng.integers(-8, 8, 200), rng.integers(0, 8, 200)], 1)
        kv = kv[np.any(kv != 0, axis=1)].astype(np.int32)
        rho = ctx.sq_rho(kv)
        sqf, sqfp, nq = ctx.sq_shells()
    orho = O.sq_rho(pos, it, 3, [L] * 3, kv)
    err = np.abs(rho - orho).max()
    print("max |drho| halfspace", err, flush=True)
    assert err < 1e-11
    assert nq.sum() == info.nvec_grid == 16 * 16 * 8 - 1
    # single-vector shell sums: total over shells = sum over all vectors
    assert sqfp.sum() > 0


@pytest.mark.parametrize("precision", ["fp64", "fp16x3", "bf16x3"])
@pytest.mark.parametrize("qset", ["ref", "halfspace"])
def test_anchor_table_bit_identical(monkeypatch, qset, precision):
    """The FP64 and 16-bit contractions with their tables of phase anchors
    (the default) and with the anchors evaluated inside the kernel
    (TRJ_SQ_ANCHOR_TABLE=0): the split partials, read back per vector
    through trj_sq_rho, are bit-identical; the shells agree to 1e-14 (their
    FP64 atomics land in a different order from run to run).  The reference
    octant set starts its anchors at 0, the half-space grid at negative kx
    and ky."""
    rng = np.random.default_rng(17)
    bsc = [1.0, 0.6, -0.3]
    if qset == "ref":
        n, box = 30000, [33.0, 33.0, 33.0]
        params = T.sq_params(3, 3, 12.0, 12.0, bsc, precision=precision, flags=T.TRJ_SQ_NO_AUTOTUNE)
        g = O.sq_setup(3, box, 12.0, 12.0)
        kv = []
        while len(kv) < 200:
            v = rng.integers(0, g.l[0] + 1, 3)
            if O.sq_member(3, g, *v) >= 0:
                kv.append(v)
    else:
        n, box = 30000, [31.0, 35.0, 29.0]
        params = T.sq_params(3, 3, 1e9, bsc=bsc, qset=T.TRJ_QSET_HALFSPACE, grid=(48, 40, 64),
                             precision=precision, flags=T.TRJ_SQ_NO_AUTOTUNE)
        kv = np.stack([rng.integers(-24, 24, 200), rng.integers(-20, 20, 200), rng.integers(1, 32, 200)], 1)
    kv = np.array(kv, dtype=np.int32)
    pos = rng.random((n, 3)) * box
    it = rng.integers(1, 4, n).astype(np.int32)
    out = {}
    for table in ("1", "0"):
        monkeypatch.setenv("TRJ_SQ_ANCHOR_TABLE", table)
        print("case: anchor table", table, qset, precision, flush=True)
        with T.Context(0) as ctx:
            ctx.sq_init(params, box)
            ctx.frame(pos, it, box, what=T.TRJ_WANT_SQ)
            sqf, sqfp, nq = ctx.sq_shells()
            out[table] = (ctx.sq_rho(kv), sqf, sqfp, nq)
    a, b = out["1"], out["0"]
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[3], b[3])
    print("shell rel diff", _rel(a[1], b[1]), _rel(a[2], b[2]), flush=True)
    assert _rel(a[1], b[1]) <= 1e-14 and _rel(a[2], b[2]) <= 1e-14
    orho = O.sq_rho(pos, it, 3, box, kv)
    err = np.abs(a[0] - orho).max()
    print("max |drho| vs oracle", err, flush=True)
    assert err < {"fp64": 1e-11, "fp16x3": 2e-3, "bf16x3": 2e-2}[precision]


def test_lattice_262k_rho_subset():
    f = next(W.lattice_262k(nframes=1))
    wl = W.LATTICE_262K
    print("case: lattice_262k rho subset", flush=True)
    with T.Context(0) as ctx:
        info = ctx.sq_init(T.sq_params(3, 3, wl.qmax, wl.qmin, wl.bsc), f.box)
        ctx.frame(f.pos, f.itype, f.box, what=T.TRJ_WANT_SQ)
        g = O.sq_setup(3, f.box, wl.qmax, wl.qmax)
        rng = np.random.default_rng(1)
        kv = []
        while len(kv) < 96:
            v = rng.integers(0, g.l[0] + 1, 3)
            if O.sq_member(3, g, *v) >= 0:
                kv.append(v)
        kv += [[1, 0, 0], [0, 1, 1], [64, 0, 0], [2, 1, 0]]
        kv = np.array(kv, dtype=np.int32)
        rho = ctx.sq_rho(kv)
        t = ctx.last_timings()
    orho = O.sq_rho(f.pos, f.itype, 3, f.box, kv)
    err = np.abs(rho - orho)
    print("contract ms", t[1], "max |drho|", err.max(), "info", info.nvec_grid, flush=True)
    # absolute 5e-12 plus 1e-15 relative (the (64,0,0) Bragg term has |rho| ~ 9e4)
    assert np.all(err < 5e-12 + 1e-15 * np.abs(orho))
    # Bragg vector (64,0,0) of the 64^3 lattice is large, (1,0,0) is tiny
    assert np.abs(orho[-2].sum()) > 1e4


@pytest.mark.parametrize("prec,tol_rho,tol_shell", [("fp32", 2e-3, 1e-5), ("fp16x3", 2e-3, 1e-5),
                                                    ("bf16x3", 2e-2, 1e-4)])
def test_reduced_precision_variants(prec, tol_rho, tol_shell):
    """FP32 / split-16-bit contraction against the oracle (per vector) and against
    the FP64 path (per shell)."""
    rng = np.random.default_rng(12)
    n, L = 30000, 33.0
    pos = rng.random((n, 3)) * L
    it = rng.integers(1, 4, n).astype(np.int32)
    kv = np.stack([rng.integers(-16, 16, 150), rng.integers(-16, 16, 150), rng.integers(0, 16, 150)], 1)
    kv = kv[np.any(kv != 0, axis=1)].astype(np.int32)
    res = {}
    print("case: precision", prec, flush=True)
    for p in ["fp64", prec]:
        with T.Context(0) as ctx:
            ctx.sq_init(T.sq_params(3, 3, 1e9, bsc=[1.0, 0.6, -0.3], qset=T.TRJ_QSET_HALFSPACE,
                                    grid=(32, 32, 32), precision=p), [L] * 3)
            ctx.frame(pos, it, [L] * 3, what=T.TRJ_WANT_SQ)
            res[p] = (ctx.sq_rho(kv), ctx.sq_shells())
    orho = O.sq_rho(pos, it, 3, [L] * 3, kv)
    err64 = np.abs(res["fp64"][0] - orho).max()
    errp = np.abs(res[prec][0] - orho).max()
    sq64, sqp = res["fp64"][1][0], res[prec][1][0]
    m = sq64 > 0
    shell_rel = np.abs(sqp[m] - sq64[m]) / sq64[m]
    print(prec, "max|drho| fp64", err64, prec, errp, "shell rel max", shell_rel.max(),
          "median", np.median(shell_rel), flush=True)
    assert err64 < 1e-10
    assert errp < tol_rho
    assert shell_rel.max() < tol_shell


def test_precision_1m_halfspace():
    """The precision study at its stated size: 1M atoms, 3 species, density 0.85,
    the 64^3 half-space grid.  Max relative error per |Q| shell against the FP64
    path: FP32 and FP16x3 within 1e-5; BF16x3 (8+8 significant bits per operand
    pair) stays within 5e-5 but is not held to 1e-5 at this size."""
    f = W.uniform_1m(seed=2)
    wl = W.UNIFORM_1M
    shells = {}
    print("case: precision 1M half-space 64^3", flush=True)
    for p in ["fp64", "fp32", "fp16x3", "bf16x3"]:
        with T.Context(0) as ctx:
            ctx.sq_init(T.sq_params(3, 3, wl.qmax, wl.qmin, wl.bsc, qset=T.TRJ_QSET_HALFSPACE, grid=wl.grid,
                                    precision=p), f.box)
            ctx.frame(f.pos, f.itype, f.box, what=T.TRJ_WANT_SQ)
            shells[p] = ctx.sq_shells()[0]
    ref = shells["fp64"]
    m = ref > 0
    worst = {}
    for p in ["fp32", "fp16x3", "bf16x3"]:
        rel = np.abs(shells[p][m] - ref[m]) / ref[m]
        worst[p] = rel.max()
        print(p, "shell rel max %.3e median %.3e" % (rel.max(), np.median(rel)), flush=True)
    assert worst["fp32"] <= 1e-5
    assert worst["fp16x3"] <= 1e-5
    assert worst["bf16x3"] <= 5e-5


@pytest.mark.parametrize("prec", ["fp64", "fp16x3", "bf16x3"])
def test_mfma_rate_probe(prec):
    """trj_mfma_rate: a dense rate between 30% and 105% of the nominal MI355X
    figure for the shape, and a plausible sustained clock."""
    nominal = {"fp64": 78.6e12, "fp16x3": 2.5e15, "bf16x3": 2.5e15}[prec]
    print("case: mfma probe", prec, flush=True)
    rate, mhz = T.mfma_rate(prec, 0)
    print(prec, "rate %.4g flop/s = %.3f of nominal, clock %.0f MHz" % (rate, rate / nominal, mhz), flush=True)
    assert 0.3 * nominal < rate < 1.05 * nominal
    assert 500.0 < mhz < 3000.0


def test_coissue_probe():
    """trj_coissue_rate: a 16x16x32 FP16 MFMA occupies its SIMD for 16 shader
    cycles (2.5e15 flop/s = 1024 SIMDs x 16384 flop / 16 cycles at 2.4 GHz);
    FMAs alone cost cycles in proportion to their count; MFMAs plus FMAs never
    take less than the MFMAs alone nor more than the two apart."""
    print("case: co-issue probe", flush=True)
    mfma = T.coissue_rate(1, "mfma")
    assert 15.0 < mfma < 20.0
    for kv in (2, 4):
        valu = T.coissue_rate(kv, "valu")
        cross = T.coissue_rate(kv, "cross")
        inter = T.coissue_rate(kv, "interleaved")
        print("valu/mfma %d: mfma %.2f valu %.2f cross %.2f interleaved %.2f cycles per MFMA"
              % (kv, mfma, valu, cross, inter), flush=True)
        assert 0.5 * kv < valu < 6.0 * kv
        for both in (cross, inter):
            assert 0.95 * mfma < both < 1.1 * (mfma + valu)
    with pytest.raises(T.TrjError):
        T.coissue_rate(5, "mfma")


def test_cluster_centres_sqfndcl():
    """TRJ_SQ_CENTRES (sqfnDcl): per frame, each octant vector's |rho|^2 / Nbcl,
    no axis vectors; frames with different point counts."""
    rng = np.random.default_rng(17)
    L = 20.0
    box = [L] * 3
    frames = [rng.random((m, 3)) * L for m in (40, 57, 33)]
    print("case: cluster centres S(Q)", flush=True)
    with T.Context(0) as ctx:
        info = ctx.sq_init(T.sq_params(3, 1, 6.0, 4.0, [1.0], flags=T.TRJ_SQ_CENTRES), box)
        for x in frames:
            ctx.frame(x, np.ones(len(x), dtype=np.int32), box, what=T.TRJ_WANT_SQ)
        sqf, _, nq = ctx.sq_shells()
    assert info.nvec_axis == 0
    g = O.sq_setup(3, box, 6.0, 4.0)
    want = np.zeros(g.nqmax + 1)
    for x in frames:
        f, _ = O.sq_frame(x, np.ones(len(x), dtype=np.int32), 3, 1, box, g, [1.0])
        # the oracle's set adds the axis shells above nqmin; keep the octant shells
        want[: g.nqmin + 1] += f[: g.nqmin + 1] / len(x)
    m = want > 0
    assert np.max(np.abs(sqf[m] - want[m]) / want[m]) < 1e-12
    assert not sqf[g.nqmin + 1:].any()


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp16x3", 1e-5)])
def test_2d_shells(precision, tol):
    """ndim = 2 (sqfnD's 2D set and sqfact2's x, y axes) against the oracle."""
    rng = np.random.default_rng(23)
    L = 40.0
    n = 3000
    pos = np.zeros((n, 3))
    pos[:, :2] = rng.random((n, 2)) * L
    it = rng.integers(1, 3, n).astype(np.int32)
    box = [L, L, 1.0]
    print("case: 2d S(Q)", precision, flush=True)
    with T.Context(0) as ctx:
        info = ctx.sq_init(T.sq_params(2, 2, 6.0, 4.5, [1.0, 0.5], precision=precision), box)
        ctx.frame(pos[:, :2].copy(), it, box, what=T.TRJ_WANT_SQ)
        sqf, sqfp, nq = ctx.sq_shells()
    g = O.sq_setup(2, box, 6.0, 4.5)
    assert info.nqmax == g.nqmax and info.nqmin == g.nqmin
    osqf, osqfp = O.sq_frame(pos, it, 2, 2, box, g, [1.0, 0.5])
    m = osqf > 0
    assert np.max(np.abs(sqf[m] - osqf[m]) / osqf[m]) <= tol
    mp = osqfp > 0
    assert np.max(np.abs(sqfp[mp] - osqfp[mp]) / osqfp[mp]) <= tol
    np.testing.assert_array_equal(nq, O.sq_nq(2, g))


def test_lattice_262k_members_and_low_shells():
    """configs[1] at its stated size: 1200 random member vectors of the
    reference q-set (plus the axis vectors' shell) against the oracle per
    vector, and the shell sums of every shell up to 6 (all their vectors
    recomputed by the oracle) within 1e-12 relative."""
    f = next(W.lattice_262k(nframes=1))
    wl = W.LATTICE_262K
    print("case: lattice_262k 1200 members + low shells", flush=True)
    g = O.sq_setup(3, f.box, wl.qmax, wl.qmax)
    rng = np.random.default_rng(2)
    kv = []
    while len(kv) < 1200:
        v = rng.integers(0, g.kext[0], 3)
        if O.sq_member(3, g, *v) >= 0:
            kv.append(v)
    low = [(a, b, c) for a in range(8) for b in range(8) for c in range(8)
           if 1 <= O.sq_member(3, g, a, b, c) <= 6]
    kv = np.array(kv + low, dtype=np.int32)
    with T.Context(0) as ctx:
        ctx.sq_init(T.sq_params(3, 3, wl.qmax, wl.qmin, wl.bsc), f.box)
        ctx.frame(f.pos, f.itype, f.box, what=T.TRJ_WANT_SQ)
        rho = ctx.sq_rho(kv)
        sqf, sqfp, _ = ctx.sq_shells()
    orho = O.sq_rho(_wrapped(f), f.itype, 3, f.box, kv)
    err = np.abs(rho - orho)
    ns = np.bincount(f.itype, minlength=4)[1:].astype(float)
    print("members: max |drho| %.3e, 99th pct %.3e, max |rho| %.3e" %
          (err.max(), np.percentile(err, 99), np.abs(orho).max()), flush=True)
    # FP64 sums of N_s unit phasors: rounding error below 4 u N_s (u = 2^-53)
    assert np.all(err < 4 * 2.0 ** -53 * ns[None, :])
    ol = orho[1200:]
    sh = np.array([O.sq_member(3, g, *v) for v in low])
    b = np.asarray(wl.bsc)
    for s in range(1, 7):
        m = sh == s
        want_p = (np.abs(ol[m]) ** 2).sum(axis=0)
        want = (np.abs(ol[m] @ b) ** 2).sum()
        rel = abs(sqf[s] - want) / want
        relp = np.max(np.abs(sqfp[:, s] - want_p) / want_p)
        print("shell", s, "vectors", int(m.sum()), "rel %.2e, partial rel %.2e" % (rel, relp), flush=True)
        assert rel <= 1e-12 and relp <= 1e-12


def test_uniform_1m_fp64_rho_against_oracle():
    """The 1M-atom FP64 path (the reference of the precision study) against the
    long-double oracle on 100 vectors of the 64^3 half-space grid."""
    f = W.uniform_1m(seed=2)
    wl = W.UNIFORM_1M
    rng = np.random.default_rng(4)
    kv = np.stack([rng.integers(-32, 32, 110), rng.integers(-32, 32, 110), rng.integers(0, 32, 110)], 1)
    kv = kv[np.any(kv != 0, axis=1)][:100].astype(np.int32)
    print("case: uniform_1m fp64 rho", flush=True)
    with T.Context(0) as ctx:
        ctx.sq_init(T.sq_params(3, 3, wl.qmax, wl.qmin, wl.bsc, qset=T.TRJ_QSET_HALFSPACE, grid=wl.grid,
                                precision="fp64"), f.box)
        ctx.frame(f.pos, f.itype, f.box, what=T.TRJ_WANT_SQ)
        rho = ctx.sq_rho(kv)
    orho = O.sq_rho(f.pos, f.itype, 3, f.box, kv)
    err = np.abs(rho - orho)
    ns = np.bincount(f.itype, minlength=4)[1:].astype(float)
    print("1M fp64: max |drho| %.3e, max |rho| %.3e" % (err.max(), np.abs(orho).max()), flush=True)
    # FP64 sums of N_s unit phasors: rounding error below 4 u N_s (u = 2^-53)
    assert np.all(err < 4 * 2.0 ** -53 * ns[None, :])
