"""Seeded synthetic frames for the benchmark workloads (numpy PCG64).

There is no trajectory data here; every workload is generated from a seed:

  lj_binary_4k      N=4000 two-species fluid, density 0.8, float32 uniform
                    positions, 10 frames from one stream seeded 0, exact 50/50
                    random species assignment; deltar=0.02, rcrdf=4, qmax=10,
                    bsc=[1.0, 0.5].
  lattice_262k      64^3 simple-cubic lattice (a = L/64, density 0.85), sites at
                    (i+1/2)a plus N(0, 0.15^2) jitter in float64, wrapped with floor;
                    species 131072/78643/52429 by a fixed random permutation;
                    jitter redrawn per frame; rcrdf=5, deltar=0.01, qmax=20,
                    bsc=[1.0, 0.6, -0.3].
  uneven_batch      frames with N log-uniform in [1e4, 1e6], density U(0.8, 0.9),
                    2-4 species with Dirichlet(1) fractions, float32 uniform
                    positions; rcrdf=4, deltar=0.01.
  uniform_1m        N=1e6 float64 uniform, density 0.85, species 0.5/0.3/0.2,
                    bsc=[1.0, 0.6, -0.3], q-grid n in [-32,31]^2 x [0,31].
  serving_stream    Poisson arrivals of bulk or z-confined slab frames, N
                    log-uniform in [1e4, 2e6], 1-4 species, float32 uniform.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Iterator, List, Optional

import numpy as np


@dataclass
class Frame:
    pos: np.ndarray        # (N, 3) float32 or float64, AoS
    itype: np.ndarray      # (N,) int32 in 1..nsp
    box: np.ndarray        # (3,) float64
    nsp: int
    periodic: tuple = (True, True, True)
    seed: int = 0


@dataclass
class Workload:
    name: str
    nsp: int
    deltar: float
    rcrdf: float
    qmax: float
    qmin: Optional[float]
    bsc: List[float]
    ndim: int = 3
    periodic: tuple = (True, True, True)
    qset: str = "ref"
    grid: tuple = (0, 0, 0)
    info: dict = field(default_factory=dict)


def _types_exact(rng, counts) -> np.ndarray:
    t = np.concatenate([np.full(c, s + 1, dtype=np.int32) for s, c in enumerate(counts)])
    rng.shuffle(t)
    return t


# ------------------------------------------------------------------ lj_binary_4k
LJ_BINARY_4K = Workload("lj_binary_4k", nsp=2, deltar=0.02, rcrdf=4.0, qmax=10.0, qmin=None,
                        bsc=[1.0, 0.5])


def lj_binary_4k(nframes: int = 10, seed: int = 0, n: int = 4000, density: float = 0.8) -> Iterator[Frame]:
    rng = np.random.Generator(np.random.PCG64(seed))
    L = (n / density) ** (1.0 / 3.0)
    box = np.array([L, L, L])
    itype = _types_exact(rng, [n // 2, n - n // 2])
    for _ in range(nframes):
        pos = (rng.random((n, 3)) * L).astype(np.float32)
        pos = np.minimum(pos, np.float32(np.nextafter(np.float32(L), np.float32(0))))
        yield Frame(pos, itype, box, 2, seed=seed)


# ------------------------------------------------------------------ lattice_262k
LATTICE_262K = Workload("lattice_262k", nsp=3, deltar=0.01, rcrdf=5.0, qmax=20.0, qmin=None,
                        bsc=[1.0, 0.6, -0.3])


def lattice_counts(n: int, fracs=(0.5, 0.3, 0.2)) -> List[int]:
    c = [int(n * f) for f in fracs[:-1]]
    c.append(n - sum(c))
    return c


def lattice_262k(nframes: int = 100, seed: int = 1, m: int = 64, density: float = 0.85,
                 sigma: float = 0.15) -> Iterator[Frame]:
    n = m ** 3
    L = (n / density) ** (1.0 / 3.0)
    a = L / m
    rng = np.random.Generator(np.random.PCG64(seed))
    g = (np.arange(m) + 0.5) * a
    sites = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3)
    itype = _types_exact(rng, lattice_counts(n))
    box = np.array([L, L, L])
    for _ in range(nframes):
        pos = sites + rng.normal(0.0, sigma, size=sites.shape)
        pos = pos - L * np.floor(pos / L)
        pos[pos >= L] -= L
        yield Frame(pos, itype, box, 3, seed=seed)


# ------------------------------------------------------------------ uneven_batch
UNEVEN_BATCH = Workload("uneven_batch", nsp=4, deltar=0.01, rcrdf=4.0, qmax=10.0, qmin=None,
                        bsc=[1.0, 1.0, 1.0, 1.0])


def uneven_frame(seed: int, nmin: float = 1e4, nmax: float = 1e6) -> Frame:
    rng = np.random.Generator(np.random.PCG64(seed))
    n = int(round(10 ** rng.uniform(np.log10(nmin), np.log10(nmax))))
    rho = rng.uniform(0.8, 0.9)
    nsp = int(rng.integers(2, 5))
    fr = rng.dirichlet(np.ones(nsp))
    counts = np.maximum(1, np.floor(fr * n).astype(int))
    counts[-1] = n - counts[:-1].sum()
    if counts[-1] < 1:
        counts = np.full(nsp, n // nsp)
        counts[-1] = n - counts[:-1].sum()
    itype = _types_exact(rng, list(counts))
    L = (n / rho) ** (1.0 / 3.0)
    pos = (rng.random((n, 3)) * L).astype(np.float32)
    return Frame(pos, itype, np.array([L, L, L]), nsp, seed=seed)


# ------------------------------------------------------------------ uniform_1m
UNIFORM_1M = Workload("uniform_1m", nsp=3, deltar=0.01, rcrdf=5.0, qmax=1.0e9, qmin=None,
                      bsc=[1.0, 0.6, -0.3], qset="halfspace", grid=(64, 64, 64))


def uniform_1m(seed: int = 2, n: int = 1_000_000, density: float = 0.85) -> Frame:
    rng = np.random.Generator(np.random.PCG64(seed))
    L = (n / density) ** (1.0 / 3.0)
    pos = rng.random((n, 3)) * L
    itype = _types_exact(rng, lattice_counts(n))
    return Frame(pos, itype, np.array([L, L, L]), 3, seed=seed)


# ------------------------------------------------------------------ serving_stream
def serving_frame(seed: int, nmin: float = 1e4, nmax: float = 2e6) -> Frame:
    rng = np.random.Generator(np.random.PCG64(seed))
    n = int(round(10 ** rng.uniform(np.log10(nmin), np.log10(nmax))))
    nsp = int(rng.integers(1, 5))
    fr = rng.dirichlet(np.ones(nsp))
    counts = np.maximum(1, np.floor(fr * n).astype(int))
    counts[-1] = n - counts[:-1].sum()
    if counts[-1] < 1:
        counts = np.full(nsp, n // nsp)
        counts[-1] = n - counts[:-1].sum()
    itype = _types_exact(rng, list(counts))
    slab = bool(rng.integers(0, 2))
    L = (n / 0.85) ** (1.0 / 3.0)
    pos = (rng.random((n, 3)) * L).astype(np.float32)
    periodic = (True, True, not slab)
    return Frame(pos, itype, np.array([L, L, L]), nsp, periodic=periodic, seed=seed)
