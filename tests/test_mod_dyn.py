"""The host side of F(Q,t): the windowed cosine transform of fftwlib.f90 on
closed forms, and the table writer."""
import numpy as np

from trj_analysis_amd import mod_dyn as M


def test_fftw1d_padding_and_grid():
    w, f = M.fftw1d(np.ones(100), 0.1, 0.0)
    # nu = int(log2(100) + 0.54) = 7, n = 256: dq = 2 pi / (256 * 0.1)
    assert len(w) == 100
    np.testing.assert_allclose(w[1], 2 * np.pi / 25.6, rtol=1e-6)


def test_fftw1d_exponential_is_lorentzian():
    tau, dt, nin = 2.0, 0.02, 2000
    t = np.arange(nin) * dt
    w, f = M.fftw1d(np.exp(-t / tau), dt, 0.0)
    # cosine transform of exp(-t/tau) over t >= 0: tau / (1 + w^2 tau^2)
    want = tau / (1 + (w * tau) ** 2)
    # trapezoid rule: an O(dt^2) absolute error
    np.testing.assert_allclose(f[1:200], want[1:200], rtol=0, atol=5e-5)
    np.testing.assert_allclose(f[0], tau, rtol=2e-3)


def test_print_fqt_tables(tmp_path):
    ndyn = 20
    t = 10.0 + 0.5 * np.arange(ndyn)
    fkt = np.vstack([np.exp(-np.arange(ndyn) / 5.0), np.exp(-np.arange(ndyn) / 3.0)])
    fskt = 0.5 * fkt
    paths = M.print_fqt(t, fkt, fskt, [1.5, 2.25], itlimit=15, tmax=0.0, outdir=str(tmp_path))
    a = np.loadtxt(paths[0])
    assert a.shape == (15, 3)
    np.testing.assert_allclose(a[:, 0], 0.5 * np.arange(15), atol=1e-3)
    np.testing.assert_allclose(a[:, 1:], fkt[:, :15].T, atol=1e-5)
    with open(paths[0]) as fh:
        assert fh.readline().startswith("#  t(tau)         F(1.5000,t)")
    s = np.loadtxt(paths[2])
    assert s.shape == (ndyn, 5)
