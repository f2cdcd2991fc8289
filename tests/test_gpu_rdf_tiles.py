"""Histograms larger than the reference allows in shared memory (rdf.cuf:61-64):
nsp = 10 (the reference's nspmax, common.cuf:41) at lsmax 501 and 801, kept
in LDS in tiles of pair-type rows, one pass per tile, with the launch shape
picked by the autotuner on the first frame.  Counts bit-exact against the
oracle's cell lists; the charge sums bit-exact through the tiled passes."""
import numpy as np
import pytest

from oracle import oracle as O
from trj_analysis_amd import trjhip as T

pytestmark = pytest.mark.gpu


def _frame(n, nsp, seed, density=0.85):
    rng = np.random.default_rng(seed)
    L = (n / density) ** (1 / 3)
    return rng.random((n, 3)) * L, rng.integers(1, nsp + 1, n).astype(np.int32), [L] * 3


@pytest.mark.parametrize("rc,autotune", [(5.0, True), (8.0, True), (5.0, False)])
def test_nsp10_tiled_histogram_bit_exact(rc, autotune):
    """lsmax 501 (55 x 501 bins, 110 KB: one pass fits the 160 KB LDS of gfx950,
    the tuner also tries 2-4 workgroups per CU with smaller tiles) and lsmax 801
    (176 KB: at least two passes)."""
    nsp, dr = 10, 0.01
    frames = [_frame(60000, nsp, s) for s in (1, 2)]
    box = frames[0][2]
    flags = 0 if autotune else T.TRJ_RDF_NO_AUTOTUNE
    print("case: nsp 10 tiles rc", rc, "autotune", autotune, flush=True)
    with T.Context(0) as ctx:
        st = ctx.rdf_init(T.rdf_params(3, nsp, dr, rc, flags=flags), box)
        for pos, it, _ in frames:
            ctx.frame(pos, it, box, what=T.TRJ_WANT_RDF)
        got = ctx.rdf_counts()
        tu = ctx.rdf_tuning()
    print("lsmax", st.lsmax, "tuning", tu, flush=True)
    assert st.nit == 55 and st.lsmax == int(round(rc / dr)) + 1
    if rc == 8.0:
        assert tu["passes"] >= 2
    if autotune:
        assert tu["tried"] >= 2
    else:
        assert tu["tried"] == 0
    want = sum(O.rdf_counts(pos, it, 3, nsp, box, (1, 1, 1), rc, dr, lsmax=st.lsmax, method="cells")
               for pos, it, _ in frames)
    np.testing.assert_array_equal(got, want)
    assert got.sum() > 0 and np.all(got.sum(axis=1) > 0)


def test_nsp10_tiles_with_charges_and_allpairs():
    """Charges through the tiled passes (the fixed-point g_qq sums are added by
    the first pass only) and a small box on the all-pairs kernel, both at
    lsmax 801 (two or more passes)."""
    nsp, dr, rc = 10, 0.01, 8.0
    pos, it, box = _frame(30000, nsp, 7)
    q = np.where(it % 2 == 0, 0.61, -0.61) + 0.03 * np.random.default_rng(3).standard_normal(len(it))
    print("case: nsp 10 tiles with charges", flush=True)
    with T.Context(0) as ctx:
        st = ctx.rdf_init(T.rdf_params(3, nsp, dr, rc, flags=T.TRJ_RDF_CHARGE), box)
        ctx.frame(pos, it, box, what=T.TRJ_WANT_RDF, charge=q)
        acc, counts = ctx.rdf_qq(), ctx.rdf_counts()
        assert ctx.rdf_tuning()["passes"] >= 2
    np.testing.assert_array_equal(
        counts, O.rdf_counts(pos, it, 3, nsp, box, (1, 1, 1), rc, dr, lsmax=st.lsmax, method="cells"))
    fixed, scale = O.rdf_qq_fixed(pos, q, 3, box, (1, 1, 1), rc, dr, lsmax=st.lsmax)
    oacc = np.zeros(st.lsmax)
    O.qq_accumulate(fixed, scale, q, oacc)
    np.testing.assert_array_equal(acc, oacc)
    # a box of fewer than 3 cells per axis: the all-pairs kernel, also tiled
    pos, it, box = _frame(4000, nsp, 8)
    print("case: nsp 10 tiles all-pairs, box", box[0], flush=True)
    with T.Context(0) as ctx:
        st = ctx.rdf_init(T.rdf_params(3, nsp, 0.005, 8.0), box)
        ctx.frame(pos, it, box, what=T.TRJ_WANT_RDF)
        got = ctx.rdf_counts()
        tu = ctx.rdf_tuning()
    print("all-pairs lsmax", st.lsmax, tu, flush=True)
    np.testing.assert_array_equal(got, O.rdf_counts(pos, it, 3, nsp, box, (1, 1, 1), 8.0, 0.005, lsmax=st.lsmax,
                                                    method="brute"))
