"""ctypes binding of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import this
module, and only as the checker.  The functions restate the reference's
algorithm; citations (reference file:line) are in trj_oracle.c.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# TRJ_ORACLE_LIB: another build of the oracle (the mutation tests load copies
# with one planted mistake each).
_LIB_PATH = os.environ.get("TRJ_ORACLE_LIB") or os.path.join(_HERE, "libtrj_oracle.so")
_lib = None


def build() -> str:
    """Compile the oracle with its Makefile (gcc); returns the library path."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


class SqGrid(C.Structure):
    _fields_ = [
        ("fk", C.c_float * 3),
        ("dq", C.c_float),
        ("nqmax", C.c_int),
        ("l", C.c_int * 3),
        ("nqmin", C.c_int),
        ("qminp", C.c_float),
        ("qmin2", C.c_float),
        ("kext", C.c_int * 3),
    ]


_D = C.POINTER(C.c_double)
_I64 = C.POINTER(C.c_int64)
_I32 = C.POINTER(C.c_int32)


def _declare(L):
    L.or_nint_d.argtypes = [C.c_double]
    L.or_nint_d.restype = C.c_long
    L.or_nint_f.argtypes = [C.c_float]
    L.or_nint_f.restype = C.c_int
    L.or_ind_ij.argtypes = [C.c_int, C.c_int, C.c_int]
    L.or_ind_ij.restype = C.c_int
    L.or_geometry.argtypes = [C.c_int, _D, _I32, _D, _D, _D, _D]
    L.or_geometry.restype = None
    L.or_wrap1.argtypes = [C.c_double, C.c_double]
    L.or_wrap1.restype = C.c_double
    L.or_rdf_lsmax.argtypes = [C.c_int, _D, C.c_double, C.c_double]
    L.or_rdf_lsmax.restype = C.c_int
    L.or_rdf_brute.argtypes = [C.c_int64, _D, _D, _D, _I32, C.c_int, C.c_int, _D,
                               C.c_double, C.c_double, C.c_int, _I64]
    L.or_rdf_brute.restype = None
    L.or_rdf_cells.argtypes = [C.c_int64, _D, _D, _D, _I32, C.c_int, C.c_int, _D, _I32,
                               _D, C.c_double, C.c_double, C.c_int, _I64]
    L.or_rdf_cells.restype = None
    L.or_rdf_gmix.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, _I64,
                              C.c_int64, _I64, _D]
    L.or_rdf_gmix.restype = None
    L.or_first_minimum.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                   C.c_int64, _I64]
    L.or_first_minimum.restype = C.c_double
    L.or_sq_setup.argtypes = [C.c_int, _D, C.c_double, C.c_double, C.POINTER(SqGrid)]
    L.or_sq_setup.restype = None
    L.or_sq_nq.argtypes = [C.c_int, C.POINTER(SqGrid), _I64]
    L.or_sq_nq.restype = None
    L.or_sq_member.argtypes = [C.c_int, C.POINTER(SqGrid), C.c_int, C.c_int, C.c_int]
    L.or_sq_member.restype = C.c_int
    L.or_sq_rho.argtypes = [C.c_int64, _D, _D, _D, _I32, C.c_int, _D, C.c_int64, _I32, _D]
    L.or_sq_rho.restype = None
    L.or_sq_frame.argtypes = [C.c_int64, _D, _D, _D, _I32, C.c_int, C.c_int, _D,
                              C.POINTER(SqGrid), _D, _D, _D]
    L.or_sq_frame.restype = None
    L.or_sq_frame_f64.argtypes = L.or_sq_frame.argtypes
    L.or_sq_frame_f64.restype = None
    L.or_sq_rho_f64.argtypes = L.or_sq_rho.argtypes
    L.or_sq_rho_f64.restype = None
    L.or_sq_columns.argtypes = [C.c_int, C.c_int, C.c_float, C.c_int64, _I64, _D, C.c_int64,
                                _D, _D, _I64, _D, _D]
    L.or_sq_columns.restype = C.c_int
    L.or_qq_scale.argtypes = [C.c_int64, _D]
    L.or_qq_scale.restype = C.c_double
    L.or_rdf_qq_brute.argtypes = [C.c_int64, _D, _D, _D, _D, C.c_int, _D, C.c_double, C.c_double,
                                  C.c_int, C.c_double, _I64]
    L.or_rdf_qq_brute.restype = None
    L.or_qq_accumulate.argtypes = [C.c_int, _I64, C.c_double, C.c_int64, _D, _D]
    L.or_qq_accumulate.restype = None
    L.or_rdf_gqq.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int64, C.c_int64, _D, _D]
    L.or_rdf_gqq.restype = None
    L.or_sq_frame_q.argtypes = [C.c_int64, _D, _D, _D, _D, C.c_int, _D, C.POINTER(SqGrid), _D]
    L.or_sq_frame_q.restype = None
    L.or_s2n.argtypes = [C.c_int64, _D, _D, C.c_int64, _D, _D, _D, C.c_int, _D, C.c_double, C.c_double, C.c_int,
                         _I64, _I64]
    L.or_s2n.restype = None
    L.or_fqt_vectors.argtypes = [C.c_int, _D, C.c_int, _D, C.c_int64, _I32, _I32]
    L.or_fqt_vectors.restype = C.c_int64
    L.or_fqt.argtypes = [C.c_int, C.c_int, _D, _D, C.c_int, _D, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int64,
                         C.c_int64, _D, _I32, _D, _D, _I64]
    L.or_fqt.restype = None
    _F = C.POINTER(C.c_float)
    L.or_slice_of.argtypes = [C.c_float, C.c_int, _F, C.c_float]
    L.or_slice_of.restype = C.c_int
    L.or_rdf_xy_brute.argtypes = [C.c_int64, _D, _D, _D, _I32, C.c_int, _D, C.c_double, C.c_double, C.c_int,
                                  C.c_int, _F, C.c_float, _I64, _I64, _I64, _D, C.c_double, _I64]
    L.or_rdf_xy_brute.restype = None
    L.or_rdf_xy_accumulate.argtypes = [C.c_int, C.c_int, C.c_int, _I64, _I64, _D]
    L.or_rdf_xy_accumulate.restype = None
    L.or_rdf_xy_gmix.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, _D, C.c_double, C.c_int64, _D, _D]
    L.or_rdf_xy_gmix.restype = None
    L.or_sq_frame_xy.argtypes = [C.c_int64, _D, _D, _D, _I32, _D, C.c_int, _D, C.POINTER(SqGrid), _D,
                                 C.c_int, _F, C.c_float, _D, _D, _D]
    L.or_sq_frame_xy.restype = None


def _p(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def nint(x: float) -> int:
    return lib().or_nint_d(float(x))


def ind_ij(i: int, j: int, nsp: int) -> int:
    return lib().or_ind_ij(i, j, nsp)


@dataclass
class Geometry:
    lmi: np.ndarray
    side2: float
    volume: float
    rcrdf: float


def geometry(ndim, box, periodic, rcrdf) -> Geometry:
    box = _f64(box)
    per = _i32(periodic)
    rc = C.c_double(rcrdf)
    lmi = np.zeros(3)
    s2 = C.c_double()
    vol = C.c_double()
    lib().or_geometry(ndim, _p(box, C.c_double), _p(per, C.c_int32), C.byref(rc),
                      _p(lmi, C.c_double), C.byref(s2), C.byref(vol))
    return Geometry(lmi, s2.value, vol.value, rc.value)


def wrap(pos: np.ndarray, box, periodic, ndim=3) -> np.ndarray:
    """pos (N,3) -> FP64 copy with periodic axes wrapped into [0,L)."""
    out = np.array(pos, dtype=np.float64, copy=True)
    w = np.vectorize(lib().or_wrap1, otypes=[np.float64])
    for k in range(ndim):
        if periodic[k]:
            out[:, k] = w(out[:, k], float(box[k]))
    return out


def rdf_lsmax(ndim, box, rcrdf, deltar) -> int:
    box = _f64(box)
    return lib().or_rdf_lsmax(ndim, _p(box, C.c_double), rcrdf, deltar)


def rdf_counts(pos, itype, ndim, nsp, box, periodic, rcrdf, deltar, lsmax=None,
               method="cells") -> np.ndarray:
    """Counts (nit, lsmax) int64 for one frame of raw positions (N,3)."""
    g = geometry(ndim, box, periodic, rcrdf)
    if lsmax is None:
        lsmax = rdf_lsmax(ndim, box, rcrdf, deltar)
    w = wrap(pos, box, periodic, ndim)
    x, y, z = (np.ascontiguousarray(w[:, k]) for k in range(3))
    it = _i32(itype)
    nit = nsp * (nsp + 1) // 2
    counts = np.zeros((nit, lsmax), dtype=np.int64)
    n = len(it)
    if method == "brute":
        lib().or_rdf_brute(n, _p(x, C.c_double), _p(y, C.c_double), _p(z, C.c_double),
                           _p(it, C.c_int32), ndim, nsp, _p(g.lmi, C.c_double), g.side2,
                           deltar, lsmax, _p(counts, C.c_int64))
    else:
        b = _f64(box)
        per = _i32(periodic)
        lib().or_rdf_cells(n, _p(x, C.c_double), _p(y, C.c_double), _p(z, C.c_double),
                           _p(it, C.c_int32), ndim, nsp, _p(b, C.c_double), _p(per, C.c_int32),
                           _p(g.lmi, C.c_double), g.side2, deltar, lsmax,
                           _p(counts, C.c_int64))
    return counts


def rdf_gmix(ndim, nsp, lsmax, deltar, volume, ntype, nconf, counts) -> np.ndarray:
    nit = nsp * (nsp + 1) // 2
    out = np.zeros((max(lsmax - 2, 0), 1 + nit))
    nt = _i64(ntype)
    c = _i64(counts)
    lib().or_rdf_gmix(ndim, nsp, lsmax, deltar, volume, _p(nt, C.c_int64), nconf,
                      _p(c, C.c_int64), _p(out, C.c_double))
    return out


def first_minimum(ndim, nsp, lsmax, deltar, volume, natoms, counts) -> float:
    c = _i64(counts)
    return lib().or_first_minimum(ndim, nsp, lsmax, deltar, volume, natoms, _p(c, C.c_int64))


def sq_setup(ndim, box, qmax, qmin) -> SqGrid:
    g = SqGrid()
    b = _f64(box)
    lib().or_sq_setup(ndim, _p(b, C.c_double), qmax, qmin, C.byref(g))
    return g


def sq_nq(ndim, g: SqGrid) -> np.ndarray:
    nq = np.zeros(g.nqmax + 1, dtype=np.int64)
    lib().or_sq_nq(ndim, C.byref(g), _p(nq, C.c_int64))
    return nq


def sq_member(ndim, g: SqGrid, kx, ky, kz) -> int:
    return lib().or_sq_member(ndim, C.byref(g), kx, ky, kz)


def sq_rho(pos, itype, nsp, box, kvecs) -> np.ndarray:
    """rho (nvec, nsp) complex128 from long-double direct sums; pos already wrapped."""
    pos = np.asarray(pos, dtype=np.float64)
    x, y, z = (np.ascontiguousarray(pos[:, k]) for k in range(3))
    it = _i32(itype)
    kv = _i32(kvecs).reshape(-1, 3)
    b = _f64(box)
    out = np.zeros((len(kv), nsp, 2))
    lib().or_sq_rho(len(it), _p(x, C.c_double), _p(y, C.c_double), _p(z, C.c_double),
                    _p(it, C.c_int32), nsp, _p(b, C.c_double), len(kv), _p(kv, C.c_int32),
                    _p(out, C.c_double))
    return out[..., 0] + 1j * out[..., 1]


def sq_rho_f64(pos, itype, nsp, box, kvecs) -> np.ndarray:
    """sq_rho with FP64 phases and OpenMP over vectors: the CPU baseline's
    timing port."""
    pos = np.asarray(pos, dtype=np.float64)
    x, y, z = (np.ascontiguousarray(pos[:, k]) for k in range(3))
    it = _i32(itype)
    kv = _i32(kvecs).reshape(-1, 3)
    b = _f64(box)
    out = np.zeros((len(kv), nsp, 2))
    lib().or_sq_rho_f64(len(it), _p(x, C.c_double), _p(y, C.c_double), _p(z, C.c_double),
                        _p(it, C.c_int32), nsp, _p(b, C.c_double), len(kv), _p(kv, C.c_int32),
                        _p(out, C.c_double))
    return out[..., 0] + 1j * out[..., 1]


def sq_frame(pos, itype, ndim, nsp, box, g: SqGrid, bsc, sqf=None, sqfp=None):
    """Accumulate one frame of the reference sum set; pos already wrapped."""
    pos = np.asarray(pos, dtype=np.float64)
    x, y, z = (np.ascontiguousarray(pos[:, k]) for k in range(3))
    it = _i32(itype)
    b = _f64(box)
    bs = _f64(bsc)
    if sqf is None:
        sqf = np.zeros(g.nqmax + 1)
    if sqfp is None:
        sqfp = np.zeros((nsp, g.nqmax + 1))
    lib().or_sq_frame(len(it), _p(x, C.c_double), _p(y, C.c_double), _p(z, C.c_double),
                      _p(it, C.c_int32), ndim, nsp, _p(b, C.c_double), C.byref(g),
                      _p(bs, C.c_double), _p(sqf, C.c_double), _p(sqfp, C.c_double))
    return sqf, sqfp


def sq_frame_f64(pos, itype, ndim, nsp, box, g: SqGrid, bsc, sqf=None, sqfp=None):
    """sq_frame with FP64 phases and OpenMP over vectors: the CPU baseline's
    timing port (same sum set and shells)."""
    pos = np.asarray(pos, dtype=np.float64)
    x, y, z = (np.ascontiguousarray(pos[:, k]) for k in range(3))
    it = _i32(itype)
    b = _f64(box)
    bs = _f64(bsc)
    if sqf is None:
        sqf = np.zeros(g.nqmax + 1)
    if sqfp is None:
        sqfp = np.zeros((nsp, g.nqmax + 1))
    lib().or_sq_frame_f64(len(it), _p(x, C.c_double), _p(y, C.c_double), _p(z, C.c_double),
                          _p(it, C.c_int32), ndim, nsp, _p(b, C.c_double), C.byref(g),
                          _p(bs, C.c_double), _p(sqf, C.c_double), _p(sqfp, C.c_double))
    return sqf, sqfp


def sq_columns(nsp, g: SqGrid, natoms, ntype, bsc, nconf, sqf, sqfp, nq):
    nt = _i64(ntype)
    bs = _f64(bsc)
    sqf = _f64(sqf)
    sqfp = _f64(sqfp)
    nq = _i64(nq)
    cols = np.zeros((g.nqmax, 7))
    mix = np.zeros((g.nqmax, 1 + nsp))
    nc = lib().or_sq_columns(nsp, g.nqmax, g.dq, natoms, _p(nt, C.c_int64), _p(bs, C.c_double),
                             nconf, _p(sqf, C.c_double), _p(sqfp, C.c_double),
                             _p(nq, C.c_int64), _p(cols, C.c_double), _p(mix, C.c_double))
    cols = cols.reshape(-1)[: g.nqmax * nc].reshape(g.nqmax, nc)
    return cols, mix


def qq_scale(q) -> float:
    q = _f64(q)
    return lib().or_qq_scale(len(q), _p(q, C.c_double))


def rdf_qq_fixed(pos, q, ndim, box, periodic, rcrdf, deltar, lsmax=None, scale=None) -> tuple:
    """(fixed int64[lsmax], scale) of one frame: sum of llrint(q_i q_j scale) per
    bin over the in-range pairs (raw positions, wrapped here)."""
    g = geometry(ndim, box, periodic, rcrdf)
    if lsmax is None:
        lsmax = rdf_lsmax(ndim, box, rcrdf, deltar)
    w = wrap(pos, box, periodic, ndim)
    x, y, z = (np.ascontiguousarray(w[:, k]) for k in range(3))
    qq = _f64(q)
    if scale is None:
        scale = qq_scale(qq)
    acc = np.zeros(lsmax, dtype=np.int64)
    lib().or_rdf_qq_brute(len(qq), _p(x, C.c_double), _p(y, C.c_double), _p(z, C.c_double),
                          _p(qq, C.c_double), ndim, _p(g.lmi, C.c_double), g.side2, deltar, lsmax,
                          scale, _p(acc, C.c_int64))
    return acc, scale


def qq_accumulate(fixed, scale, q, acc):
    """acc += (fixed/scale)/avq2 in place (avq2 = sum q^2 / N)."""
    f = _i64(fixed)
    qq = _f64(q)
    lib().or_qq_accumulate(len(f), _p(f, C.c_int64), scale, len(qq), _p(qq, C.c_double),
                           _p(acc, C.c_double))
    return acc


def rdf_gqq(ndim, lsmax, deltar, volume, natoms, nconf, acc) -> np.ndarray:
    a = _f64(acc)
    out = np.zeros(max(lsmax - 2, 0))
    lib().or_rdf_gqq(ndim, lsmax, deltar, volume, natoms, nconf, _p(a, C.c_double), _p(out, C.c_double))
    return out


def sq_frame_q(pos, q, ndim, box, g: SqGrid, sqfq=None):
    """Accumulate the charge-weighted sums of the reference sum set; pos wrapped."""
    pos = np.asarray(pos, dtype=np.float64)
    x, y, z = (np.ascontiguousarray(pos[:, k]) for k in range(3))
    qq = _f64(q)
    b = _f64(box)
    if sqfq is None:
        sqfq = np.zeros(g.nqmax + 1)
    lib().or_sq_frame_q(len(qq), _p(x, C.c_double), _p(y, C.c_double), _p(z, C.c_double),
                        _p(qq, C.c_double), ndim, _p(b, C.c_double), C.byref(g), _p(sqfq, C.c_double))
    return sqfq


# ------------------------------------------------------------------ slices
def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def slice_of(z, zslice, zgrid) -> int:
    zs = _f32(zslice)
    return lib().or_slice_of(float(np.float32(z)), len(zs), _p(zs, C.c_float), float(np.float32(zgrid)))


def rdf_xy_counts(pos, itype, nsp, box, rcrdf, deltar, zslice, zgrid, lsmax=None, q=None, scale=None):
    """One frame of rdfxy: (counts (nslice, nit, lsmax) int64, count (nslice,),
    countp (nslice, nsp), fixed (nslice, lsmax) or None, scale).  pos (N,3) as
    ingested (x, y wrapped; z as given); box is the frame's box (Lz enters the
    cutoff cap as in 3D)."""
    pos = np.asarray(pos, dtype=np.float64)
    g = geometry(3, box, (1, 1, 0), rcrdf)
    if lsmax is None:
        lsmax = rdf_lsmax(3, box, rcrdf, deltar)
    zs = _f32(zslice)
    ns = len(zs)
    nit = nsp * (nsp + 1) // 2
    x, y, z = (np.ascontiguousarray(pos[:, k]) for k in range(3))
    it = _i32(itype)
    counts = np.zeros((ns, nit, lsmax), dtype=np.int64)
    count = np.zeros(ns, dtype=np.int64)
    countp = np.zeros((ns, nsp), dtype=np.int64)
    lxy = _f64([box[0], box[1]])
    fixed = None
    qq = None
    if q is not None:
        qq = _f64(q)
        if scale is None:
            scale = qq_scale(qq)
        fixed = np.zeros((ns, lsmax), dtype=np.int64)
    lib().or_rdf_xy_brute(len(it), _p(x, C.c_double), _p(y, C.c_double), _p(z, C.c_double), _p(it, C.c_int32),
                          nsp, _p(lxy, C.c_double), g.side2, deltar, lsmax, ns, _p(zs, C.c_float),
                          float(np.float32(zgrid)), _p(counts, C.c_int64), _p(count, C.c_int64),
                          _p(countp, C.c_int64), _p(qq, C.c_double) if qq is not None else None,
                          float(scale) if scale is not None else 0.0,
                          _p(fixed, C.c_int64) if fixed is not None else None)
    return counts, count, countp, fixed, scale


def rdf_xy_accumulate(nsp, counts, countp, hxy=None):
    counts = _i64(counts)
    countp = _i64(countp)
    ns, nit, ls = counts.shape
    if hxy is None:
        hxy = np.zeros((ns, nsp, nsp, ls))
    lib().or_rdf_xy_accumulate(nsp, ls, ns, _p(counts, C.c_int64), _p(countp, C.c_int64), _p(hxy, C.c_double))
    return hxy


def rdf_xy_gmix(nsp, lsmax, deltar, box, zgrid, nconf, hxy):
    """(npairs, lsmax-2, nslice) g_xy tables, pairs in ind_ij order."""
    h = _f64(hxy)
    ns = h.shape[0]
    nit = nsp * (nsp + 1) // 2
    out = np.zeros((nit, max(lsmax - 2, 0), ns))
    b = _f64(box)
    lib().or_rdf_xy_gmix(nsp, lsmax, ns, deltar, _p(b, C.c_double), float(np.float32(zgrid)), nconf,
                         _p(h, C.c_double), _p(out, C.c_double))
    return out


def sq_frame_xy(pos, itype, nsp, box, g: SqGrid, bsc, zslice, zgrid, q=None, acc=None):
    """Accumulate one frame of sqf3xy/sqf2xy; returns (sqf, sqfp, sqfq) with
    shapes (nslice, nqmax+1), (nslice, nsp, nqmax+1), (nslice, nqmax+1)."""
    pos = np.asarray(pos, dtype=np.float64)
    x, y, z = (np.ascontiguousarray(pos[:, k]) for k in range(3))
    it = _i32(itype)
    zs = _f32(zslice)
    ns = len(zs)
    if acc is None:
        acc = (np.zeros((ns, g.nqmax + 1)), np.zeros((ns, nsp, g.nqmax + 1)), np.zeros((ns, g.nqmax + 1)))
    sqf, sqfp, sqfq = acc
    b = _f64(box)
    bs = _f64(bsc)
    qq = _f64(q) if q is not None else None
    lib().or_sq_frame_xy(len(it), _p(x, C.c_double), _p(y, C.c_double), _p(z, C.c_double), _p(it, C.c_int32),
                         _p(qq, C.c_double) if qq is not None else None, nsp, _p(b, C.c_double), C.byref(g),
                         _p(bs, C.c_double), ns, _p(zs, C.c_float), float(np.float32(zgrid)),
                         _p(sqf, C.c_double), _p(sqfp, C.c_double), _p(sqfq, C.c_double))
    return acc


def s2n_sums(pts, pos, ndim, box, periodic, rcrdf, deltar, lsmax=None, gi=None, g2i=None):
    """Add one frame of the number-fluctuation sums for unit points pts (np, 3);
    pos raw (N, 3), wrapped here.  Returns (gi, g2i) int64 (lsmax+1,)."""
    g = geometry(ndim, box, periodic, rcrdf)
    if lsmax is None:
        lsmax = rdf_lsmax(ndim, box, rcrdf, deltar)
    w = wrap(pos, box, periodic, ndim)
    x, y, z = (np.ascontiguousarray(w[:, k]) for k in range(3))
    pt = _f64(pts)
    b = _f64(box)
    if gi is None:
        gi = np.zeros(lsmax + 1, dtype=np.int64)
        g2i = np.zeros(lsmax + 1, dtype=np.int64)
    lib().or_s2n(len(pt), _p(pt, C.c_double), _p(b, C.c_double), len(x), _p(x, C.c_double), _p(y, C.c_double),
                 _p(z, C.c_double), ndim, _p(g.lmi, C.c_double), g.side2, deltar, lsmax, _p(gi, C.c_int64),
                 _p(g2i, C.c_int64))
    return gi, g2i


def fqt_vectors(ndim, box, qw):
    """(kv (m,3) int32, bin (m,)) of the S(Q,w) q-set, grouped by qw."""
    b = _f64(box)
    q = _f64(qw)
    m = lib().or_fqt_vectors(ndim, _p(b, C.c_double), len(q), _p(q, C.c_double), 0, None, None)
    kv = np.zeros((max(m, 1), 3), dtype=np.int32)
    bn = np.zeros(max(m, 1), dtype=np.int32)
    lib().or_fqt_vectors(ndim, _p(b, C.c_double), len(q), _p(q, C.c_double), m, _p(kv, C.c_int32), _p(bn, C.c_int32))
    return kv[:m], bn[:m]


def fqt(frames, itype, ndim, nsp, bsc, box, qw, nbuffer, jump, ndyn, itlimit):
    """(tot (nqw, ndyn), tots (nqw, ndyn), count_org) over frames (list of (N,3) positions)."""
    pos = np.ascontiguousarray(np.stack([np.asarray(f, dtype=np.float64) for f in frames]))
    it = _i32(itype)
    b = _f64(box)
    q = _f64(qw)
    bs = _f64(bsc)
    tot = np.zeros((len(q), ndyn))
    tots = np.zeros((len(q), ndyn))
    co = C.c_int64(0)
    lib().or_fqt(ndim, nsp, _p(bs, C.c_double), _p(b, C.c_double), len(q), _p(q, C.c_double), nbuffer, jump,
                 ndyn, itlimit, pos.shape[0], pos.shape[1], _p(pos, C.c_double), _p(it, C.c_int32),
                 _p(tot, C.c_double), _p(tots, C.c_double), C.byref(co))
    return tot, tots, co.value

