"""The host side of F(Q,t): the fkt.dat / fskt.dat table writer."""
import numpy as np

from trj_analysis_amd import mod_dyn as M


def test_ndyn_of():
    # dyn_init (dynamics.cuf:96): nconf - (nbuffer - 1) * jump lags
    assert M.ndyn_of(100, 4, 10) == 70


def test_print_fqt_tables(tmp_path):
    ndyn = 20
    t = 10.0 + 0.5 * np.arange(ndyn)
    fkt = np.vstack([np.exp(-np.arange(ndyn) / 5.0), np.exp(-np.arange(ndyn) / 3.0)])
    fskt = 0.5 * fkt
    paths = M.print_fqt(t, fkt, fskt, [1.5, 2.25], itlimit=15, outdir=str(tmp_path))
    assert len(paths) == 2
    a = np.loadtxt(paths[0])
    assert a.shape == (15, 3)
    np.testing.assert_allclose(a[:, 0], 0.5 * np.arange(15), atol=1e-3)
    np.testing.assert_allclose(a[:, 1:], fkt[:, :15].T, atol=1e-5)
    with open(paths[0]) as fh:
        assert fh.readline().startswith("#  t(tau)         F(1.5000,t)")
    b = np.loadtxt(paths[1])
    np.testing.assert_allclose(b[:, 1:], fskt[:, :15].T, atol=1e-5)
    with open(paths[1]) as fh:
        assert fh.readline().startswith("#  t(tau)         Fs(1.5000,t)")
