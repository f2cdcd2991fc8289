"""The freud-tool replacements on the GPU: g(r) and S(k) of a small NetCDF
file against direct numpy evaluations of the restated definitions (pairs by
brute force, phase sums over the same vectors), and the files they write."""
import os

import numpy as np
import pytest

from trj_analysis_amd.tools import fast_grfreud_netcdf as G
from trj_analysis_amd.tools import fast_sq_freud_netcdf as S

pytestmark = pytest.mark.gpu
scipy_io = pytest.importorskip("scipy.io")


def _write(path, coords, cell):
    nf, na, _ = coords.shape
    with scipy_io.netcdf_file(path, "w", version=2) as nc:
        nc.createDimension("frame", None)
        nc.createDimension("atom", na)
        nc.createDimension("spatial", 3)
        nc.createDimension("cell_spatial", 3)
        cl = nc.createVariable("cell_lengths", "f", ("frame", "cell_spatial"))
        ty = nc.createVariable("type", "i", ("frame", "atom"))
        xyz = nc.createVariable("coordinates", "f", ("frame", "atom", "spatial"))
        for f in range(nf):
            cl[f] = cell
            ty[f] = np.ones(na, dtype=np.int32)
            xyz[f] = coords[f]


def _file(tmp_path, nf=2, na=1200, L=18.0, seed=3):
    rng = np.random.default_rng(seed)
    coords = (rng.random((nf, na, 3)) * L).astype(np.float32)
    path = str(tmp_path / "t.nc")
    _write(path, coords, np.array([L, L, L], dtype=np.float32))
    return path, coords, L


def test_gr_against_brute_force(tmp_path):
    path, coords, L = _file(tmp_path)
    bins, rmax = 40, 6.0
    print("case: tools gr", flush=True)
    r, g, nf = G.compute(path, bins, rmax, say=lambda *a: None)
    dr = rmax / bins
    hist = np.zeros(bins)
    for f in range(nf):
        x = coords[f].astype(np.float64)
        d = x[None] - x[:, None]
        d -= L * np.round(d / L)
        rr = np.sqrt((d ** 2).sum(-1))[~np.eye(len(x), dtype=bool)]
        h, _ = np.histogram(rr[rr < rmax], bins=np.arange(bins + 1) * dr)
        hist += h
    edges = np.arange(bins + 1) * dr
    shell = 4.0 / 3.0 * np.pi * (edges[1:] ** 3 - edges[:-1] ** 3)
    n = coords.shape[1]
    want = hist / (n * (n / L ** 3) * nf * shell)
    np.testing.assert_allclose(g, want.astype(np.float32), rtol=2e-6)
    np.testing.assert_allclose(r, (edges[:-1] + dr / 2).astype(np.float32))
    assert abs(np.mean(g[bins // 2:]) - 1.0) < 0.05          # ideal gas
    out = tmp_path / "gr.dat"
    assert G.main([path, "--bins", "40", "--r-max", "6", "--out", str(out)]) == 0
    with open(out) as fh:
        assert fh.readline().strip() == "# r g(r)"
    np.testing.assert_allclose(np.loadtxt(out), np.column_stack((r, g)), rtol=1e-6)


def test_sq_against_direct_sums(tmp_path):
    path, coords, L = _file(tmp_path, na=800, L=30.0, seed=4)
    bins, kmax = 20, 1.2
    print("case: tools sq", flush=True)
    k, s, nf = S.compute(path, bins, kmax, say=lambda *a: None)
    kmin = 2 * np.pi / L
    kv, kk, _ = S._vectors([L] * 3, kmin, kmax)
    idx = np.floor((kk - kmin) / ((kmax - kmin) / bins)).astype(int)
    sums = np.zeros(bins)
    cnt = np.zeros(bins)
    for f in range(nf):
        x = coords[f].astype(np.float64)
        ph = 2 * np.pi * (x @ (kv.T / L))
        rho2 = np.abs(np.exp(1j * ph).sum(0)) ** 2 / len(x)
        np.add.at(sums, idx, rho2)
        np.add.at(cnt, idx, 1.0)
    m = cnt > 0
    want = np.zeros(bins)
    want[m] = sums[m] / cnt[m]
    np.testing.assert_allclose(s[m], want[m].astype(np.float32), rtol=2e-6)
    assert np.all(np.isnan(s[~m]))
    # every +-k pair once: the half-space set has no vector whose negative is also in it
    tup = {tuple(v) for v in kv}
    assert not any(tuple(-v) in tup for v in kv)
