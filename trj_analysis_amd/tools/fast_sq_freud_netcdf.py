"""S(k) of a NetCDF trajectory, as tools/fast_sq_freud_netcdf.py computes it
with freud (freud.diffraction.StaticStructureFactorDirect, bins=60,
k_min = 2 pi / max(L), k_max = 0.5, reset=False; fast_sq_freud_netcdf.py:
9-10, 23-41), on the GPU phase sums.

    python -m trj_analysis_amd.tools.fast_sq_freud_netcdf traj.nc [--bins 60] [--k-max 0.5]

writes Sq.dat ("k S(k)", numpy.savetxt, as fast_sq_freud_netcdf.py:49).

The direct structure factor, restated (freud is not installed here): over the
reciprocal-lattice vectors k of the first frame's box with k_min <= |k| <
k_max, S(k_i) is the mean of |sum_j exp(i k.r_j)|^2 / N over the vectors of
bin i (width (k_max - k_min)/bins) and over frames; k is the bin centre, an
empty bin gives nan.  One vector of each +-k pair is summed (S(k) = S(-k)).
freud subsamples bins holding more than 10000 vectors at random; that is not
reproduced (every vector is summed).  |rho(k)| comes from the library's
contraction on a half-space q-grid, FP64.
"""
from __future__ import annotations

import argparse
import os
import sys

import numpy as np

from .. import trjhip as T


def _vectors(box, kmin, kmax):
    """One vector of each +-k pair with kmin <= |k| < kmax: (kv int32 (m,3), |k| (m,))."""
    fk = 2.0 * np.pi / np.asarray(box, dtype=np.float64)
    nmax = [int(np.floor(kmax / f)) for f in fk]
    ax = [np.arange(-m, m + 1) for m in nmax]
    gx, gy, gz = np.meshgrid(ax[0], ax[1], np.arange(0, nmax[2] + 1), indexing="ij")
    kv = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], 1)
    x, y, z = kv[:, 0], kv[:, 1], kv[:, 2]
    half = (z > 0) | ((z == 0) & (y > 0)) | ((z == 0) & (y == 0) & (x > 0))
    kv = kv[half]
    kk = np.sqrt(((kv * fk) ** 2).sum(1))
    keep = (kk >= kmin) & (kk < kmax)
    return kv[keep].astype(np.int32), kk[keep], nmax


def compute(path: str, bins: int = 60, k_max: float = 0.5, device: int = 0, say=print):
    with T.NcTrajectory(path) as tr:
        nframes = tr.nframes
        with T.Context(device) as ctx:
            sums = np.zeros(bins)
            cnt = np.zeros(bins)
            k_min = None
            for frame in range(nframes):
                pos, _, _, cell = tr.read(frame)
                box = [float(c) for c in cell]
                if frame == 0:
                    k_min = 2.0 * np.pi / max(box)
                    kv, kk, nmax = _vectors(box, k_min, k_max)
                    dk = (k_max - k_min) / bins
                    idx = np.floor((kk - k_min) / dk).astype(int)
                    ok = (idx >= 0) & (idx < bins)
                    kv, idx = kv[ok], idx[ok]
                    grid = tuple(2 * (m + 1) for m in nmax)
                    dq = min(2.0 * np.pi / b for b in box)
                    ctx.sq_init(T.sq_params(3, 1, k_max + 2 * dq, qset=T.TRJ_QSET_HALFSPACE, grid=grid), box)
                p32 = np.ascontiguousarray(pos, dtype=np.float32)
                n = len(p32)
                ctx.frame(p32, np.ones(n, dtype=np.int32), box, what=T.TRJ_WANT_SQ)
                if len(kv):
                    rho = ctx.sq_rho(kv)[:, 0]
                    np.add.at(sums, idx, np.abs(rho) ** 2 / n)
                    np.add.at(cnt, idx, 1.0)
                if frame % 10 == 0:
                    say("processed frame", frame)
    with np.errstate(invalid="ignore", divide="ignore"):
        s = sums / cnt
    k = k_min + (np.arange(bins) + 0.5) * (k_max - k_min) / bins
    return k.astype(np.float32), s.astype(np.float32), nframes


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("trajfile")
    ap.add_argument("--bins", type=int, default=60)
    ap.add_argument("--k-max", type=float, default=0.5)
    ap.add_argument("--device", type=int, default=int(os.environ.get("TRJ_DEVICE", "0")))
    ap.add_argument("--out", default="Sq.dat")
    a = ap.parse_args(argv)
    k, s, nf = compute(a.trajfile, a.bins, a.k_max, a.device)
    np.savetxt(a.out, np.column_stack((k, s)), header="k S(k)")
    print("Finished. Frames processed:", nf)
    return 0


if __name__ == "__main__":
    sys.exit(main())
