"""ctypes binding of libtrjhip.so (the C ABI declared in include/trj_hip.h).

The native library is the only implementation: importing this module never
falls back to a CPU path, and any call that needs the GPU raises TrjError
when the library is missing or no device is available.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtrjhip.so")

TRJ_OK = 0
TRJ_EINVAL, TRJ_EDEVICE, TRJ_ENOMEM, TRJ_ESTATE, TRJ_ERANGE, TRJ_ELIMIT = -1, -2, -3, -4, -5, -6
TRJ_F32, TRJ_F64 = 0, 1
TRJ_AOS, TRJ_SOA = 0, 1
TRJ_WANT_RDF, TRJ_WANT_SQ, TRJ_WANT_FQT = 1, 2, 4
TRJ_QSET_REF, TRJ_QSET_HALFSPACE = 0, 1
TRJ_PREC_FP64, TRJ_PREC_FP32, TRJ_PREC_FP16X3, TRJ_PREC_BF16X3 = 0, 1, 2, 3
TRJ_SQ_NQ_CONSISTENT, TRJ_SQ_CHARGE, TRJ_SQ_CENTRES, TRJ_SQ_NO_AUTOTUNE = 1, 2, 4, 8
TRJ_RDF_CHARGE, TRJ_RDF_FLOOR_BINS, TRJ_RDF_NO_AUTOTUNE = 1, 2, 4
TRJ_PIPE_SERIAL, TRJ_PIPE_OVERLAP = 0, 1
TRJ_STATE_REPLACE, TRJ_STATE_ADD = 0, 1

PRECISIONS = {"fp64": TRJ_PREC_FP64, "fp32": TRJ_PREC_FP32, "fp16x3": TRJ_PREC_FP16X3,
              "bf16x3": TRJ_PREC_BF16X3}

# Every symbol the header declares (checked by the CPU test suite).
EXPORTS = [
    "trj_abi_version", "trj_create", "trj_destroy", "trj_last_error", "trj_rdf_init",
    "trj_sq_init", "trj_sq_plan", "trj_rdf_plan", "trj_frame", "trj_frame_device", "trj_frame_device_q", "trj_frame_device_noalloc", "trj_rdf_tuning", "trj_rdf_cell_grid", "trj_sq_tuning", "trj_sync",
    "trj_rdf_counts", "trj_rdf_gmix", "trj_rdf_first_minimum", "trj_sq_shells", "trj_sq_columns",
    "trj_sq_rho", "trj_fqt_init", "trj_fqt_results", "trj_rdf_sampling", "trj_rdf_sampling_points", "trj_rdf_s2n_sums", "trj_rdf_s2n",
    "trj_set_slices", "trj_slice_counts", "trj_rdf_xy", "trj_rdf_xy_gmix", "trj_rdf_xy_qq",
    "trj_rdf_xy_gcc", "trj_sq_xy", "trj_sq_xy_columns", "trj_frame_q", "trj_rdf_qq", "trj_rdf_gqq", "trj_sq_qq", "trj_sq_qq_column", "trj_reset", "trj_last_timings", "trj_timing_totals", "trj_timing_reset",
    "trj_device_bytes", "trj_set_pipeline", "trj_state_bytes", "trj_state_save", "trj_state_load",
    "trj_lds_atomic_rate", "trj_mfma_rate", "trj_coissue_rate", "trj_nc_open", "trj_nc_close", "trj_nc_last_error", "trj_nc_info", "trj_nc_read",
    "trj_nc_time", "trj_nc_frame_ref", "trj_nc_frame_ref_q", "trj_nc_keep_nonperiodic",
    "trj_batch_create", "trj_batch_destroy", "trj_batch_last_error",
    "trj_batch_layout", "trj_batch_run", "trj_batch_autotune", "trj_batch_sq_layout", "trj_batch_sq_run",
    "trj_srv_create", "trj_srv_destroy",
    "trj_srv_last_error", "trj_request_footprint", "trj_srv_submit", "trj_srv_poll", "trj_srv_result",
    "trj_srv_stats", "trj_srv_memory",
]


class TrjError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"trj_hip error {code}: {msg}")
        self.code = code


class RdfParams(C.Structure):
    _fields_ = [("ndim", C.c_int32), ("nsp", C.c_int32), ("deltar", C.c_double),
                ("rcrdf", C.c_double), ("periodic", C.c_int32 * 3), ("flags", C.c_uint32)]


class SqParams(C.Structure):
    _fields_ = [("ndim", C.c_int32), ("nsp", C.c_int32), ("qmax", C.c_double),
                ("qmin", C.c_double), ("bsc", C.c_double * 10), ("qset", C.c_int32),
                ("grid", C.c_int32 * 3), ("precision", C.c_int32), ("flags", C.c_uint32)]


class FqtParams(C.Structure):
    _fields_ = [("ndim", C.c_int32), ("nsp", C.c_int32), ("nqw", C.c_int32), ("nbuffer", C.c_int32),
                ("jump", C.c_int32), ("ndyn", C.c_int32), ("itlimit", C.c_int32), ("pad", C.c_int32),
                ("qw", C.POINTER(C.c_double)), ("bsc", C.POINTER(C.c_double))]


class SqInfo(C.Structure):
    _fields_ = [("dq", C.c_double), ("fk", C.c_double * 3), ("qmin_eff", C.c_double),
                ("nqmax", C.c_int32), ("nqmin", C.c_int32), ("lmax", C.c_int32 * 3),
                ("nvec_grid", C.c_int64), ("nvec_axis", C.c_int64), ("box_ext", C.c_int64 * 3)]


class BatchFrame(C.Structure):
    _fields_ = [("n", C.c_int64), ("nsp", C.c_int32), ("periodic", C.c_int32 * 3),
                ("box", C.c_double * 3)]


class BatchParams(C.Structure):
    _fields_ = [("ndim", C.c_int32), ("deltar", C.c_double), ("rcrdf", C.c_double),
                ("frames_per_launch", C.c_int32), ("waves_per_cell", C.c_int32),
                ("workgroups", C.c_int32), ("use_graph", C.c_int32)]


class BatchSqParams(C.Structure):
    _fields_ = [("grid", C.c_int32 * 3), ("precision", C.c_int32), ("qmax", C.c_double)]


class Request(C.Structure):
    _fields_ = [("n", C.c_int64), ("pos", C.c_void_p), ("dtype", C.c_int32), ("itype", C.c_void_p),
                ("box", C.c_double * 3), ("what", C.c_uint32), ("rdf", RdfParams), ("sq", SqParams)]


_lib = None


def lib():
    """Load libtrjhip.so (built by trj_analysis_amd.build); raise if absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise TrjError(TRJ_EDEVICE, f"{LIB_PATH} is not built; run trj_analysis_amd.build.build()")
        L = C.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


def _declare(L):
    vp, d, i32, i64, u32 = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int32), \
        C.POINTER(C.c_int64), C.c_uint32
    L.trj_abi_version.restype = C.c_int
    L.trj_create.argtypes = [C.POINTER(vp), C.c_int]
    L.trj_destroy.argtypes = [vp]
    L.trj_destroy.restype = None
    L.trj_last_error.argtypes = [vp]
    L.trj_last_error.restype = C.c_char_p
    L.trj_rdf_init.argtypes = [vp, C.POINTER(RdfParams), d, i32, i32]
    L.trj_sq_init.argtypes = [vp, C.POINTER(SqParams), d, C.POINTER(SqInfo)]
    L.trj_sq_plan.argtypes = [C.POINTER(SqParams), d, C.POINTER(SqInfo), i64]
    L.trj_rdf_plan.argtypes = [C.POINTER(RdfParams), d, i32, d]
    L.trj_frame.argtypes = [vp, vp, C.c_int32, C.c_int32, i32, C.c_int64, d, u32]
    L.trj_frame_q.argtypes = [vp, vp, C.c_int32, C.c_int32, i32, d, C.c_int64, d, u32]
    L.trj_rdf_qq.argtypes = [vp, d]
    L.trj_set_slices.argtypes = [vp, C.c_int32, C.POINTER(C.c_float), C.c_float]
    L.trj_rdf_sampling.argtypes = [vp, C.c_int32, C.c_uint64]
    L.trj_fqt_init.argtypes = [vp, C.POINTER(FqtParams), d, C.POINTER(C.c_int32), i64]
    L.trj_fqt_results.argtypes = [vp, d, d, i64]
    L.trj_rdf_sampling_points.argtypes = [vp, d]
    L.trj_rdf_s2n_sums.argtypes = [vp, i64, i64]
    L.trj_rdf_s2n.argtypes = [vp, C.c_int64, d]
    L.trj_slice_counts.argtypes = [vp, i64, i64]
    L.trj_rdf_xy.argtypes = [vp, d, i64]
    L.trj_rdf_xy_gmix.argtypes = [vp, C.c_int64, d]
    L.trj_rdf_xy_qq.argtypes = [vp, d]
    L.trj_rdf_xy_gcc.argtypes = [vp, C.c_int64, d]
    L.trj_sq_xy.argtypes = [vp, d, d, d]
    L.trj_sq_xy_columns.argtypes = [vp, C.c_int64, d, d, d]
    L.trj_rdf_gqq.argtypes = [vp, C.c_int64, d]
    L.trj_sq_qq.argtypes = [vp, d]
    L.trj_sq_qq_column.argtypes = [vp, C.c_int64, d]
    L.trj_frame_device.argtypes = [vp, vp, C.c_int32, C.c_int32, vp, C.c_int64, d, u32]
    L.trj_frame_device_q.argtypes = [vp, vp, C.c_int32, C.c_int32, vp, vp, C.c_int64, d, u32]
    L.trj_frame_device_noalloc.argtypes = [vp, vp, C.c_int32, C.c_int32, vp, C.c_int64, d, u32]
    L.trj_rdf_tuning.argtypes = [vp, i32, i32, i32, d, i32]
    L.trj_rdf_cell_grid.argtypes = [vp, i32, d]
    L.trj_sq_tuning.argtypes = [vp, i32, i32, i32, d, i32]
    L.trj_sync.argtypes = [vp]
    L.trj_rdf_counts.argtypes = [vp, i64]
    L.trj_rdf_gmix.argtypes = [vp, C.c_int64, d]
    L.trj_rdf_first_minimum.argtypes = [vp, d]
    L.trj_sq_shells.argtypes = [vp, d, d, i64]
    L.trj_sq_columns.argtypes = [vp, C.c_int64, d, i32, d]
    L.trj_sq_rho.argtypes = [vp, C.c_int64, i32, d]
    L.trj_reset.argtypes = [vp]
    L.trj_last_timings.argtypes = [vp, d]
    L.trj_timing_totals.argtypes = [vp, d, i64]
    L.trj_timing_reset.argtypes = [vp]
    L.trj_device_bytes.argtypes = [vp]
    L.trj_device_bytes.restype = C.c_int64
    L.trj_set_pipeline.argtypes = [vp, C.c_int]
    L.trj_state_bytes.argtypes = [vp]
    L.trj_state_bytes.restype = C.c_int64
    L.trj_state_save.argtypes = [vp, vp, C.c_int64]
    L.trj_state_load.argtypes = [vp, vp, C.c_int64, C.c_int]
    L.trj_lds_atomic_rate.argtypes = [C.c_int, d]
    L.trj_nc_open.argtypes = [C.POINTER(vp), C.c_char_p, C.c_char_p, C.c_int32]
    L.trj_nc_close.argtypes = [vp]
    L.trj_nc_close.restype = None
    L.trj_nc_last_error.argtypes = [vp]
    L.trj_nc_last_error.restype = C.c_char_p
    L.trj_nc_info.argtypes = [vp, C.POINTER(NcFileInfo)]
    L.trj_nc_read.argtypes = [vp, C.c_int64, d, i32, d, d]
    L.trj_mfma_rate.argtypes = [C.c_int, C.c_int, d, d]
    L.trj_coissue_rate.argtypes = [C.c_int, C.c_int, C.c_int, d]
    L.trj_nc_time.argtypes = [vp, C.c_int64, d]
    L.trj_nc_frame_ref.argtypes = [vp, C.c_int64, C.c_int32, i32, d, i32, i64, d, i32]
    L.trj_nc_frame_ref_q.argtypes = [vp, C.c_int64, C.c_int32, i32, d, i32, d, i64, d, i32]
    L.trj_nc_keep_nonperiodic.argtypes = [vp, C.c_int32]
    L.trj_batch_create.argtypes = [C.POINTER(vp), C.c_int, C.POINTER(BatchParams)]
    L.trj_batch_destroy.argtypes = [vp]
    L.trj_batch_destroy.restype = None
    L.trj_batch_last_error.argtypes = [vp]
    L.trj_batch_last_error.restype = C.c_char_p
    L.trj_batch_layout.argtypes = [C.POINTER(BatchParams), C.c_int32, C.POINTER(BatchFrame), i32, i32,
                                   i64, i64]
    L.trj_batch_run.argtypes = [vp, C.c_int32, C.POINTER(BatchFrame), vp, C.c_int32, vp, i64, d]
    L.trj_batch_sq_layout.argtypes = [C.POINTER(BatchParams), C.POINTER(BatchSqParams), C.c_int32,
                                      C.POINTER(BatchFrame), i32, i64, i64, i64, i64]
    L.trj_batch_sq_run.argtypes = [vp, C.POINTER(BatchSqParams), C.c_int32, C.POINTER(BatchFrame), vp, C.c_int32,
                                   vp, d, d, d, i64, d]
    L.trj_batch_autotune.argtypes = [vp, C.c_int32, C.POINTER(BatchFrame), vp, C.c_int32, vp,
                                     C.POINTER(BatchParams)]
    L.trj_srv_create.argtypes = [C.POINTER(vp), C.c_int, C.c_int32, C.c_int64]
    L.trj_srv_destroy.argtypes = [vp]
    L.trj_srv_destroy.restype = None
    L.trj_srv_last_error.argtypes = [vp]
    L.trj_srv_last_error.restype = C.c_char_p
    L.trj_request_footprint.argtypes = [C.POINTER(Request)]
    L.trj_request_footprint.restype = C.c_int64
    L.trj_srv_submit.argtypes = [vp, C.POINTER(Request), i64]
    L.trj_srv_poll.argtypes = [vp, C.c_int64, i32]
    L.trj_srv_result.argtypes = [vp, C.c_int64, i64, d, d, i64]
    L.trj_srv_stats.argtypes = [vp, i32, i32, i64]
    L.trj_srv_memory.argtypes = [vp, i64, i64, i64]
    no_int = ("trj_destroy", "trj_last_error", "trj_device_bytes", "trj_batch_destroy",
              "trj_batch_last_error", "trj_srv_destroy", "trj_srv_last_error", "trj_request_footprint",
              "trj_state_bytes", "trj_nc_close", "trj_nc_last_error")
    for name in EXPORTS:
        if name not in no_int:
            getattr(L, name).restype = C.c_int


class NcFileInfo(C.Structure):
    _fields_ = [("nframes", C.c_int64), ("natoms", C.c_int64), ("version", C.c_int32),
                ("coord_dtype", C.c_int32), ("has_type", C.c_int32), ("has_cell_origin", C.c_int32),
                ("has_cell_lengths", C.c_int32), ("has_time", C.c_int32), ("has_charge", C.c_int32)]


class NcTrajectory:
    """A LAMMPS NetCDF trajectory (classic formats) through trj_nc_*: frames as
    stored (read) or as the reference's read_nc_cfg hands them on (frame_ref)."""

    def __init__(self, path: str):
        self._L = lib()
        h = C.c_void_p()
        err = C.create_string_buffer(512)
        rc = self._L.trj_nc_open(C.byref(h), os.fsencode(path), err, 512)
        if rc != TRJ_OK:
            raise TrjError(rc, f"trj_nc_open: {err.value.decode()}")
        self._h = h
        info = NcFileInfo()
        self._check(self._L.trj_nc_info(self._h, C.byref(info)), "trj_nc_info")
        self.nframes, self.natoms, self.version = info.nframes, info.natoms, info.version
        self.coord_dtype = info.coord_dtype
        self.has_type = bool(info.has_type)
        self.has_cell_origin = bool(info.has_cell_origin)
        self.has_cell_lengths = bool(info.has_cell_lengths)
        self.has_time = bool(info.has_time)
        self.has_charge = bool(info.has_charge)

    def _check(self, rc, what):
        if rc != TRJ_OK:
            raise TrjError(rc, f"{what}: {self._L.trj_nc_last_error(self._h).decode()}")

    def read(self, frame: int):
        pos = np.zeros((self.natoms, 3))
        typ = np.zeros(self.natoms, dtype=np.int32)
        org, cell = np.zeros(3), np.zeros(3)
        self._check(self._L.trj_nc_read(self._h, int(frame), _ptr(pos, C.c_double),
                                        _ptr(typ, C.c_int32) if self.has_type else None,
                                        _ptr(org, C.c_double), _ptr(cell, C.c_double)), "trj_nc_read")
        return pos, (typ if self.has_type else None), org, cell

    def keep_nonperiodic(self, keep: bool = True):
        """Hand non-periodic coordinates on as read (trj_nc_keep_nonperiodic)."""
        self._check(self._L.trj_nc_keep_nonperiodic(self._h, int(bool(keep))), "trj_nc_keep_nonperiodic")

    def time(self, frame: int) -> float:
        """time(frame) as stored (the reference's nstep is its integer part)."""
        t = C.c_double(0.0)
        self._check(self._L.trj_nc_time(self._h, int(frame), C.byref(t)), "trj_nc_time")
        return t.value

    def frame_ref(self, frame: int, sp_types, periodic=(True, True, True), charges: bool = False):
        """(pos SoA (3, n), itype (n,), box (3,), periodic (3,)) as read_nc_cfg hands
        them on; periodic in is the /INPUT/ flags, out has zero-length axes cleared.
        With charges=True a fifth item, the atoms' charges q (n,), is returned
        (trj_nc_frame_ref_q)."""
        st = np.ascontiguousarray(np.asarray(sp_types, dtype=np.int32))
        pos = np.empty(3 * self.natoms)
        it = np.empty(self.natoms, dtype=np.int32)
        n = C.c_int64()
        box = np.zeros(3)
        per = np.array([1 if p else 0 for p in periodic][:3], dtype=np.int32)
        if charges:
            q = np.empty(self.natoms)
            self._check(self._L.trj_nc_frame_ref_q(self._h, int(frame), len(st), _ptr(st, C.c_int32),
                                                   _ptr(pos, C.c_double), _ptr(it, C.c_int32),
                                                   _ptr(q, C.c_double), C.byref(n), _ptr(box, C.c_double),
                                                   _ptr(per, C.c_int32)), "trj_nc_frame_ref_q")
        else:
            self._check(self._L.trj_nc_frame_ref(self._h, int(frame), len(st), _ptr(st, C.c_int32),
                                                 _ptr(pos, C.c_double), _ptr(it, C.c_int32), C.byref(n),
                                                 _ptr(box, C.c_double), _ptr(per, C.c_int32)), "trj_nc_frame_ref")
        m = n.value
        if charges:
            return pos[:3 * m].reshape(3, m), it[:m], box, per, q[:m]
        return pos[:3 * m].reshape(3, m), it[:m], box, per

    def close(self):
        if getattr(self, "_h", None):
            self._L.trj_nc_close(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        self.close()


def lds_atomic_rate(device: int = 0) -> float:
    """LDS-atomic adds per second of the device (trj_lds_atomic_rate)."""
    out = C.c_double(0.0)
    rc = lib().trj_lds_atomic_rate(int(device), C.byref(out))
    if rc != TRJ_OK:
        raise TrjError(rc, "trj_lds_atomic_rate")
    return out.value


def mfma_rate(precision="bf16x3", device: int = 0) -> tuple[float, float]:
    """(dense flops per second, sustained shader clock in MHz) of one MFMA shape
    (trj_mfma_rate): "fp64", "fp16x3" or "bf16x3"."""
    p = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
    f, mhz = C.c_double(0.0), C.c_double(0.0)
    rc = lib().trj_mfma_rate(int(device), p, C.byref(f), C.byref(mhz))
    if rc != TRJ_OK:
        raise TrjError(rc, "trj_mfma_rate")
    return f.value, mhz.value


COISSUE_MODES = {"mfma": 0, "valu": 1, "cross": 2, "interleaved": 3}


def coissue_rate(valu_per_mfma: int, mode="cross", device: int = 0) -> float:
    """Shader cycles per 16-bit MFMA of one SIMD with valu_per_mfma FP32 FMAs
    per MFMA (trj_coissue_rate): mode "mfma" (MFMAs only), "valu" (FMAs only),
    "cross" (one wavefront of each per SIMD) or "interleaved" (both in every
    wavefront)."""
    m = COISSUE_MODES[mode] if isinstance(mode, str) else int(mode)
    c = C.c_double(0.0)
    rc = lib().trj_coissue_rate(int(device), int(valu_per_mfma), m, C.byref(c))
    if rc != TRJ_OK:
        raise TrjError(rc, "trj_coissue_rate")
    return c.value


def _ptr(a: np.ndarray, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def _box3(box) -> np.ndarray:
    b = np.ones(3, dtype=np.float64)
    bb = np.asarray(box, dtype=np.float64).ravel()
    b[: len(bb)] = bb
    return b


def rdf_params(ndim, nsp, deltar, rcrdf, periodic=(True, True, True), flags=0) -> RdfParams:
    p = RdfParams()
    p.ndim, p.nsp, p.deltar, p.rcrdf = ndim, nsp, float(deltar), float(rcrdf)
    for k in range(3):
        p.periodic[k] = int(bool(periodic[k]))
    p.flags = flags
    return p


def sq_params(ndim, nsp, qmax, qmin=None, bsc=None, qset=TRJ_QSET_REF, grid=(0, 0, 0),
              precision="fp64", flags=0) -> SqParams:
    p = SqParams()
    p.ndim, p.nsp, p.qmax = ndim, nsp, float(qmax)
    p.qmin = float(qmin) if qmin is not None else -1.0
    b = [1.0] * nsp if bsc is None else list(bsc)
    for s in range(10):
        p.bsc[s] = float(b[s]) if s < len(b) else 1.0
    p.qset = qset
    for k in range(3):
        p.grid[k] = int(grid[k])
    p.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
    p.flags = flags
    return p


def sq_plan(p: SqParams, box) -> tuple[SqInfo, np.ndarray]:
    """Host-only q-set plan (no GPU): info and shell counts nq(0:nqmax)."""
    info = SqInfo()
    b = _box3(box)
    rc = lib().trj_sq_plan(C.byref(p), _ptr(b, C.c_double), C.byref(info), None)
    if rc != TRJ_OK:
        raise TrjError(rc, "trj_sq_plan failed")
    nq = np.zeros(info.nqmax + 1, dtype=np.int64)
    lib().trj_sq_plan(C.byref(p), _ptr(b, C.c_double), C.byref(info), _ptr(nq, C.c_int64))
    return info, nq


def rdf_plan(p: RdfParams, box) -> tuple[int, float]:
    b = _box3(box)
    ls = C.c_int32()
    rc = C.c_double()
    r = lib().trj_rdf_plan(C.byref(p), _ptr(b, C.c_double), C.byref(ls), C.byref(rc))
    if r != TRJ_OK:
        raise TrjError(r, "trj_rdf_plan failed")
    return ls.value, rc.value


@dataclass
class RdfState:
    lsmax: int
    nit: int
    nsp: int
    ndim: int


class Context:
    """One trj_hip context (one HIP stream on one device)."""

    def __init__(self, device: int = 0):
        self._L = lib()
        h = C.c_void_p()
        rc = self._L.trj_create(C.byref(h), device)
        if rc != TRJ_OK:
            raise TrjError(rc, f"trj_create(device={device}) failed (no HIP device?)")
        self._h = h
        self.rdf: Optional[RdfState] = None
        self.sq_info: Optional[SqInfo] = None
        self.sq_nsp = 0
        self.nslice = 0
        self.nrandom = 0

    # -- lifetime
    def close(self):
        if getattr(self, "_h", None):
            self._L.trj_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc: int, what: str):
        if rc != TRJ_OK:
            msg = self._L.trj_last_error(self._h)
            raise TrjError(rc, f"{what}: {msg.decode() if msg else ''}")

    # -- init
    def rdf_init(self, p: RdfParams, box) -> RdfState:
        b = _box3(box)
        ls, nit = C.c_int32(), C.c_int32()
        self._check(self._L.trj_rdf_init(self._h, C.byref(p), _ptr(b, C.c_double), C.byref(ls),
                                         C.byref(nit)), "trj_rdf_init")
        self.rdf = RdfState(ls.value, nit.value, p.nsp, p.ndim)
        return self.rdf

    def sq_init(self, p: SqParams, box) -> SqInfo:
        b = _box3(box)
        info = SqInfo()
        self._check(self._L.trj_sq_init(self._h, C.byref(p), _ptr(b, C.c_double), C.byref(info)),
                    "trj_sq_init")
        self.sq_info = info
        self.sq_nsp = p.nsp
        return info

    # -- frames
    def frame(self, pos: np.ndarray, itype: np.ndarray, box, what=TRJ_WANT_RDF | TRJ_WANT_SQ,
              layout=TRJ_AOS, charge=None):
        """One frame (trj_frame, or trj_frame_q when per-atom charges are given)."""
        pos = np.ascontiguousarray(pos)
        if pos.dtype not in (np.float32, np.float64):
            pos = pos.astype(np.float64)
        dtype = TRJ_F32 if pos.dtype == np.float32 else TRJ_F64
        it = np.ascontiguousarray(itype, dtype=np.int32)
        n = len(it)
        b = _box3(box)
        if charge is None:
            self._check(self._L.trj_frame(self._h, pos.ctypes.data_as(C.c_void_p), dtype, layout,
                                          _ptr(it, C.c_int32), n, _ptr(b, C.c_double), what),
                        "trj_frame")
        else:
            q = np.ascontiguousarray(charge, dtype=np.float64)
            if len(q) != n:
                raise TrjError(TRJ_EINVAL, "charge must have one value per atom")
            self._check(self._L.trj_frame_q(self._h, pos.ctypes.data_as(C.c_void_p), dtype, layout,
                                            _ptr(it, C.c_int32), _ptr(q, C.c_double), n,
                                            _ptr(b, C.c_double), what), "trj_frame_q")

    def frame_device(self, pos_ptr: int, dtype: int, layout: int, itype_ptr: int, n: int, box,
                     what=TRJ_WANT_RDF | TRJ_WANT_SQ, charge_ptr: Optional[int] = None):
        """Device-resident frame (trj_frame_device, or trj_frame_device_q with
        charge_ptr: n FP64 charges in device memory)."""
        b = _box3(box)
        if charge_ptr is None:
            self._check(self._L.trj_frame_device(self._h, C.c_void_p(pos_ptr), dtype, layout,
                                                 C.c_void_p(itype_ptr), n, _ptr(b, C.c_double), what),
                        "trj_frame_device")
        else:
            self._check(self._L.trj_frame_device_q(self._h, C.c_void_p(pos_ptr), dtype, layout,
                                                   C.c_void_p(itype_ptr), C.c_void_p(charge_ptr), n,
                                                   _ptr(b, C.c_double), what), "trj_frame_device_q")

    def rdf_tuning(self) -> dict:
        """The RDF autotuner's choice (trj_rdf_tuning)."""
        r, p, w, t = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        ms = C.c_double()
        self._check(self._L.trj_rdf_tuning(self._h, C.byref(r), C.byref(p), C.byref(w), C.byref(ms), C.byref(t)),
                    "trj_rdf_tuning")
        return {"rows_per_pass": r.value, "passes": p.value, "waves_per_cell": w.value, "ms": ms.value,
                "tried": t.value}

    def rdf_cell_grid(self) -> dict:
        """Cells per axis of the last RDF frame and the cell side over the
        cutoff (trj_rdf_cell_grid)."""
        nc = (C.c_int32 * 3)()
        f = C.c_double()
        self._check(self._L.trj_rdf_cell_grid(self._h, nc, C.byref(f)), "trj_rdf_cell_grid")
        return {"cells": [nc[0], nc[1], nc[2]], "side_over_cutoff": f.value}

    def sq_tuning(self) -> dict:
        """The S(Q) autotuner's choice (trj_sq_tuning)."""
        k, a, x, t = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        ms = C.c_double()
        self._check(self._L.trj_sq_tuning(self._h, C.byref(k), C.byref(a), C.byref(x), C.byref(ms), C.byref(t)),
                    "trj_sq_tuning")
        return {"splits": k.value, "atoms_per_split": a.value, "xcd_grouped": x.value, "ms": ms.value,
                "tried": t.value}

    def frame_device_noalloc(self, pos_ptr: int, dtype: int, layout: int, itype_ptr: int, n: int, box):
        """RDF of a device-resident frame without any allocation
        (trj_frame_device_noalloc; per-frame buffers from the library's
        static scratch)."""
        b = _box3(box)
        self._check(self._L.trj_frame_device_noalloc(self._h, C.c_void_p(pos_ptr), dtype, layout,
                                                     C.c_void_p(itype_ptr), n, _ptr(b, C.c_double), TRJ_WANT_RDF),
                    "trj_frame_device_noalloc")

    def sync(self):
        self._check(self._L.trj_sync(self._h), "trj_sync")

    # -- results
    def rdf_counts(self) -> np.ndarray:
        r = self.rdf
        out = np.zeros((r.nit, r.lsmax), dtype=np.int64)
        self._check(self._L.trj_rdf_counts(self._h, _ptr(out, C.c_int64)), "trj_rdf_counts")
        return out

    def rdf_gmix(self, nconf: int = 0) -> np.ndarray:
        r = self.rdf
        out = np.zeros((max(r.lsmax - 2, 0), 1 + r.nit))
        self._check(self._L.trj_rdf_gmix(self._h, nconf, _ptr(out, C.c_double)), "trj_rdf_gmix")
        return out

    # ------------------------------------------------------------ F(Q,t)
    def fqt_init(self, qw, nbuffer: int, jump: int, ndyn: int, itlimit: int, box, ndim: int = 3, nsp: int = 1,
                 bsc=None):
        """F(Q,t) over time origins (trj_fqt_init): returns (nvec, nw)."""
        q = np.ascontiguousarray(np.asarray(qw, dtype=np.float64).ravel())
        b = np.ascontiguousarray(np.asarray([1.0] * nsp if bsc is None else bsc, dtype=np.float64))
        p = FqtParams(ndim, nsp, len(q), nbuffer, jump, ndyn, itlimit, 0, _ptr(q, C.c_double), _ptr(b, C.c_double))
        nv = C.c_int32(0)
        nw = np.zeros(len(q), dtype=np.int64)
        bx = _box3(box)
        self._check(self._L.trj_fqt_init(self._h, C.byref(p), _ptr(bx, C.c_double), C.byref(nv), _ptr(nw, C.c_int64)),
                    "trj_fqt_init")
        self.fqt_shape = (len(q), ndyn)
        return nv.value, nw

    def fqt_results(self):
        """(fkt (nqw, ndyn), fskt (nqw, ndyn), count_org), normalised as print_rtcor."""
        f = np.zeros(self.fqt_shape)
        fs = np.zeros(self.fqt_shape)
        co = C.c_int64(0)
        self._check(self._L.trj_fqt_results(self._h, _ptr(f, C.c_double), _ptr(fs, C.c_double), C.byref(co)),
                    "trj_fqt_results")
        return f, fs, co.value

    # ------------------------------------------------ number fluctuations
    def rdf_sampling(self, nrandom: int, seed: int = 12345):
        """Fixed random points for the number fluctuations (trj_rdf_sampling)."""
        self._check(self._L.trj_rdf_sampling(self._h, int(nrandom), int(seed)), "trj_rdf_sampling")
        self.nrandom = int(nrandom)

    def rdf_sampling_points(self) -> np.ndarray:
        out = np.zeros((self.nrandom, 3))
        self._check(self._L.trj_rdf_sampling_points(self._h, _ptr(out, C.c_double)), "trj_rdf_sampling_points")
        return out

    def rdf_s2n_sums(self):
        gi = np.zeros(self.rdf.lsmax + 1, dtype=np.int64)
        g2i = np.zeros(self.rdf.lsmax + 1, dtype=np.int64)
        self._check(self._L.trj_rdf_s2n_sums(self._h, _ptr(gi, C.c_int64), _ptr(g2i, C.c_int64)), "trj_rdf_s2n_sums")
        return gi, g2i

    def rdf_s2n(self, nconf: int = 0) -> np.ndarray:
        out = np.zeros(max(self.rdf.lsmax - 2, 0))
        self._check(self._L.trj_rdf_s2n(self._h, nconf, _ptr(out, C.c_double)), "trj_rdf_s2n")
        return out

    # ------------------------------------------------------------ slices
    def set_slices(self, zslice, zgrid):
        """Confined slices (trj_set_slices); before rdf_init / sq_init.  An
        empty zslice turns them off."""
        z = np.ascontiguousarray(np.asarray(zslice, dtype=np.float32).ravel())
        self._check(self._L.trj_set_slices(self._h, len(z), z.ctypes.data_as(C.POINTER(C.c_float)),
                                           float(np.float32(zgrid))), "trj_set_slices")
        self.nslice = len(z)
        self.zslice = z.copy()
        self.zgrid = np.float32(zgrid)

    def slice_counts(self):
        """(count (nslice,), countp (nslice, nsp)) of the last frame (rdfxy rule)."""
        nsp = self.rdf.nsp if self.rdf else self.sq_nsp
        a = np.zeros(self.nslice, dtype=np.int64)
        b = np.zeros((self.nslice, nsp), dtype=np.int64)
        self._check(self._L.trj_slice_counts(self._h, _ptr(a, C.c_int64), _ptr(b, C.c_int64)), "trj_slice_counts")
        return a, b

    def rdf_xy(self):
        """(hxy (nslice, nit, lsmax) normalised per frame, raw counts (nslice, nit, lsmax))."""
        h = np.zeros((self.nslice, self.rdf.nit, self.rdf.lsmax))
        c = np.zeros((self.nslice, self.rdf.nit, self.rdf.lsmax), dtype=np.int64)
        self._check(self._L.trj_rdf_xy(self._h, _ptr(h, C.c_double), _ptr(c, C.c_int64)), "trj_rdf_xy")
        return h, c

    def rdf_xy_gmix(self, nconf: int = 0) -> np.ndarray:
        """gxy tables (nit, lsmax-2, nslice)."""
        out = np.zeros((self.rdf.nit, max(self.rdf.lsmax - 2, 0), self.nslice))
        self._check(self._L.trj_rdf_xy_gmix(self._h, nconf, _ptr(out, C.c_double)), "trj_rdf_xy_gmix")
        return out

    def rdf_xy_qq(self) -> np.ndarray:
        out = np.zeros((self.nslice, self.rdf.lsmax))
        self._check(self._L.trj_rdf_xy_qq(self._h, _ptr(out, C.c_double)), "trj_rdf_xy_qq")
        return out

    def rdf_xy_gcc(self, nconf: int = 0) -> np.ndarray:
        """gxy_cc table (lsmax-2, nslice)."""
        out = np.zeros((max(self.rdf.lsmax - 2, 0), self.nslice))
        self._check(self._L.trj_rdf_xy_gcc(self._h, nconf, _ptr(out, C.c_double)), "trj_rdf_xy_gcc")
        return out

    def sq_xy(self, charges: bool = False):
        """(sqf (nslice, nqmax+1), sqfp (nslice, nsp, nqmax+1), sqfq or None)."""
        ns = self.sq_info.nqmax + 1
        f = np.zeros((self.nslice, ns))
        fp = np.zeros((self.nslice, self.sq_nsp, ns))
        fq = np.zeros((self.nslice, ns)) if charges else None
        self._check(self._L.trj_sq_xy(self._h, _ptr(f, C.c_double), _ptr(fp, C.c_double),
                                      _ptr(fq, C.c_double) if charges else None), "trj_sq_xy")
        return f, fp, fq

    def sq_xy_columns(self, nconf: int = 0, charges: bool = False):
        """(sqxy (nqmax, 1+nslice), sqpxy (nsp, nqmax, 1+nslice), sqqxy or None)."""
        nq = self.sq_info.nqmax
        a = np.zeros((nq, 1 + self.nslice))
        b = np.zeros((self.sq_nsp, nq, 1 + self.nslice))
        q = np.zeros((nq, 1 + self.nslice)) if charges else None
        self._check(self._L.trj_sq_xy_columns(self._h, nconf, _ptr(a, C.c_double), _ptr(b, C.c_double),
                                              _ptr(q, C.c_double) if charges else None), "trj_sq_xy_columns")
        return a, b, q

    def rdf_qq(self) -> np.ndarray:
        """Charge-weighted accumulator gqq[k-1], k = 1..lsmax (trj_rdf_qq)."""
        out = np.zeros(self.rdf.lsmax)
        self._check(self._L.trj_rdf_qq(self._h, _ptr(out, C.c_double)), "trj_rdf_qq")
        return out

    def rdf_gqq(self, nconf: int = 0) -> np.ndarray:
        """printrdf's g_qq column for r_i, i = 1..lsmax-2 (trj_rdf_gqq)."""
        out = np.zeros(max(self.rdf.lsmax - 2, 0))
        self._check(self._L.trj_rdf_gqq(self._h, nconf, _ptr(out, C.c_double)), "trj_rdf_gqq")
        return out

    def rdf_first_minimum(self) -> float:
        v = C.c_double()
        self._check(self._L.trj_rdf_first_minimum(self._h, C.byref(v)), "trj_rdf_first_minimum")
        return v.value

    def sq_shells(self):
        ns = self.sq_info.nqmax + 1
        sqf = np.zeros(ns)
        sqfp = np.zeros((self.sq_nsp, ns))
        nq = np.zeros(ns, dtype=np.int64)
        self._check(self._L.trj_sq_shells(self._h, _ptr(sqf, C.c_double), _ptr(sqfp, C.c_double),
                                          _ptr(nq, C.c_int64)), "trj_sq_shells")
        return sqf, sqfp, nq

    def sq_columns(self, nconf: int = 0):
        nq = self.sq_info.nqmax
        cols = np.zeros((nq, 7))
        mix = np.zeros((nq, 1 + self.sq_nsp))
        nc = C.c_int32()
        self._check(self._L.trj_sq_columns(self._h, nconf, _ptr(cols, C.c_double), C.byref(nc),
                                           _ptr(mix, C.c_double)), "trj_sq_columns")
        cols = cols.reshape(-1)[: nq * nc.value].reshape(nq, nc.value)
        return cols, mix

    def sq_qq(self) -> np.ndarray:
        """Charge-weighted shell sums sqfq[0..nqmax] (trj_sq_qq)."""
        out = np.zeros(self.sq_info.nqmax + 1)
        self._check(self._L.trj_sq_qq(self._h, _ptr(out, C.c_double)), "trj_sq_qq")
        return out

    def sq_qq_column(self, nconf: int = 0) -> np.ndarray:
        """printSQ's S_qq column for shells 1..nqmax (trj_sq_qq_column)."""
        out = np.zeros(self.sq_info.nqmax)
        self._check(self._L.trj_sq_qq_column(self._h, nconf, _ptr(out, C.c_double)), "trj_sq_qq_column")
        return out

    def sq_rho(self, kvecs: Sequence) -> np.ndarray:
        kv = np.ascontiguousarray(np.asarray(kvecs, dtype=np.int32).reshape(-1, 3))
        out = np.zeros((len(kv), self.sq_nsp, 2))
        self._check(self._L.trj_sq_rho(self._h, len(kv), _ptr(kv, C.c_int32), _ptr(out, C.c_double)),
                    "trj_sq_rho")
        return out[..., 0] + 1j * out[..., 1]

    def reset(self):
        self._check(self._L.trj_reset(self._h), "trj_reset")

    def last_timings(self) -> np.ndarray:
        out = np.zeros(4)
        self._check(self._L.trj_last_timings(self._h, _ptr(out, C.c_double)), "trj_last_timings")
        return out

    def timing_totals(self) -> tuple[np.ndarray, int]:
        out = np.zeros(4)
        nf = C.c_int64()
        self._check(self._L.trj_timing_totals(self._h, _ptr(out, C.c_double), C.byref(nf)),
                    "trj_timing_totals")
        return out, nf.value

    def timing_reset(self):
        self._check(self._L.trj_timing_reset(self._h), "trj_timing_reset")

    def device_bytes(self) -> int:
        return self._L.trj_device_bytes(self._h)

    def state_save(self) -> bytes:
        """The context's accumulated results as bytes (trj_state_save)."""
        nb = self._L.trj_state_bytes(self._h)
        if nb < 0:
            raise TrjError(nb, "trj_state_bytes")
        buf = C.create_string_buffer(nb)
        self._check(self._L.trj_state_save(self._h, buf, nb), "trj_state_save")
        return buf.raw

    def state_load(self, blob: bytes, mode: int = TRJ_STATE_REPLACE):
        """Replace (TRJ_STATE_REPLACE) or add to (TRJ_STATE_ADD) the accumulators."""
        buf = C.create_string_buffer(bytes(blob), len(blob))
        self._check(self._L.trj_state_load(self._h, buf, len(blob), int(mode)), "trj_state_load")

    def set_pipeline(self, mode: int):
        """TRJ_PIPE_SERIAL or TRJ_PIPE_OVERLAP (RDF on its own stream beside S(Q))."""
        self._check(self._L.trj_set_pipeline(self._h, int(mode)), "trj_set_pipeline")


def batch_params(ndim, deltar, rcrdf, frames_per_launch=0, waves_per_cell=0, workgroups=0,
                 use_graph=True) -> BatchParams:
    p = BatchParams()
    p.ndim, p.deltar, p.rcrdf = ndim, float(deltar), float(rcrdf)
    p.frames_per_launch, p.waves_per_cell, p.workgroups = frames_per_launch, waves_per_cell, workgroups
    p.use_graph = int(bool(use_graph))
    return p


def batch_frames(frames) -> "C.Array":
    """frames: iterable of (n, nsp, box, periodic)."""
    frames = list(frames)
    arr = (BatchFrame * len(frames))()
    for k, (n, nsp, box, per) in enumerate(frames):
        arr[k].n, arr[k].nsp = int(n), int(nsp)
        b = _box3(box)
        for a in range(3):
            arr[k].box[a] = float(b[a])
            arr[k].periodic[a] = int(bool(per[a]))
    return arr


def batch_layout(p: BatchParams, frames_arr):
    nf = len(frames_arr)
    ls = np.zeros(nf, dtype=np.int32)
    nit = np.zeros(nf, dtype=np.int32)
    off = np.zeros(nf, dtype=np.int64)
    tot = C.c_int64()
    r = lib().trj_batch_layout(C.byref(p), nf, frames_arr, _ptr(ls, C.c_int32), _ptr(nit, C.c_int32),
                               _ptr(off, C.c_int64), C.byref(tot))
    if r != TRJ_OK:
        raise TrjError(r, "trj_batch_layout failed")
    return ls, nit, off, tot.value


def batch_sq_params(grid=(64, 64, 64), precision=TRJ_PREC_FP64, qmax=1.0e9) -> BatchSqParams:
    q = BatchSqParams()
    for a in range(3):
        q.grid[a] = int(grid[a])
    q.precision, q.qmax = int(precision), float(qmax)
    return q


def batch_sq_layout(p: BatchParams, q: BatchSqParams, frames_arr):
    """nqmax, sqf/nq offsets, sqfp offsets and the two totals of trj_batch_sq_layout."""
    nf = len(frames_arr)
    nqmax = np.zeros(nf, dtype=np.int32)
    off = np.zeros(nf, dtype=np.int64)
    offp = np.zeros(nf, dtype=np.int64)
    t1, t2 = C.c_int64(), C.c_int64()
    r = lib().trj_batch_sq_layout(C.byref(p), C.byref(q), nf, frames_arr, _ptr(nqmax, C.c_int32),
                                  _ptr(off, C.c_int64), _ptr(offp, C.c_int64), C.byref(t1), C.byref(t2))
    if r != TRJ_OK:
        raise TrjError(r, "trj_batch_sq_layout failed")
    return nqmax, off, offp, t1.value, t2.value


class Batch:
    """Batched frames of uneven size (trj_batch_*): per-frame pair histograms
    (run) and per-frame S(Q) shell sums (sq_run)."""

    def __init__(self, p: BatchParams, device: int = 0):
        self._L = lib()
        h = C.c_void_p()
        rc = self._L.trj_batch_create(C.byref(h), device, C.byref(p))
        if rc != TRJ_OK:
            raise TrjError(rc, "trj_batch_create failed")
        self._h = h
        self.p = p

    def close(self):
        if getattr(self, "_h", None):
            self._L.trj_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != TRJ_OK:
            msg = self._L.trj_batch_last_error(self._h)
            raise TrjError(rc, f"{what}: {msg.decode() if msg else ''}")

    def run(self, frames_arr, pos_ptr: int, dtype: int, itype_ptr: int, want_hist=True):
        ls, nit, off, tot = batch_layout(self.p, frames_arr)
        hist = np.zeros(max(tot, 1), dtype=np.int64) if want_hist else None
        ms = C.c_double()
        self._check(self._L.trj_batch_run(self._h, len(frames_arr), frames_arr, C.c_void_p(pos_ptr), dtype,
                                          C.c_void_p(itype_ptr),
                                          _ptr(hist, C.c_int64) if want_hist else None, C.byref(ms)),
                    "trj_batch_run")
        out = None
        if want_hist:
            out = [hist[off[k]:off[k] + nit[k] * ls[k]].reshape(nit[k], ls[k]) for k in range(len(off))]
        return out, ms.value

    def sq_run(self, q: BatchSqParams, frames_arr, pos_ptr: int, dtype: int, itype_ptr: int, bsc=None):
        """trj_batch_sq_run: per frame (sqf, sqfp [nsp][nqmax+1], nq), and the device ms.
        bsc: per-frame scattering lengths (list of sequences), None for all 1."""
        nqmax, off, offp, t1, t2 = batch_sq_layout(self.p, q, frames_arr)
        nf = len(frames_arr)
        sqf = np.zeros(max(t1, 1))
        sqfp = np.zeros(max(t2, 1))
        nq = np.zeros(max(t1, 1), dtype=np.int64)
        b = None
        if bsc is not None:
            b = np.zeros((nf, 10))
            for k, row in enumerate(bsc):
                b[k, :len(row)] = row
        ms = C.c_double()
        self._check(self._L.trj_batch_sq_run(self._h, C.byref(q), nf, frames_arr, C.c_void_p(pos_ptr), dtype,
                                             C.c_void_p(itype_ptr), _ptr(b, C.c_double) if b is not None else None,
                                             _ptr(sqf, C.c_double), _ptr(sqfp, C.c_double), _ptr(nq, C.c_int64),
                                             C.byref(ms)), "trj_batch_sq_run")
        out = []
        for k in range(nf):
            ns = int(nqmax[k]) + 1
            nsp = frames_arr[k].nsp
            out.append((sqf[off[k]:off[k] + ns].copy(), sqfp[offp[k]:offp[k] + nsp * ns].reshape(nsp, ns).copy(),
                        nq[off[k]:off[k] + ns].copy()))
        return out, ms.value

    def autotune(self, frames_arr, pos_ptr: int, dtype: int, itype_ptr: int) -> BatchParams:
        chosen = BatchParams()
        self._check(self._L.trj_batch_autotune(self._h, len(frames_arr), frames_arr, C.c_void_p(pos_ptr),
                                               dtype, C.c_void_p(itype_ptr), C.byref(chosen)),
                    "trj_batch_autotune")
        self.p = chosen
        return chosen


def request(pos: np.ndarray, itype: np.ndarray, box, what=TRJ_WANT_RDF | TRJ_WANT_SQ, rdf=None,
            sq=None) -> "tuple[Request, tuple]":
    """A trj_request over host arrays; returns (request, keepalive)."""
    pos = np.ascontiguousarray(pos)
    if pos.dtype not in (np.float32, np.float64):
        pos = pos.astype(np.float64)
    it = np.ascontiguousarray(itype, dtype=np.int32)
    r = Request()
    r.n = len(it)
    r.pos = pos.ctypes.data
    r.dtype = TRJ_F32 if pos.dtype == np.float32 else TRJ_F64
    r.itype = it.ctypes.data
    b = _box3(box)
    for k in range(3):
        r.box[k] = float(b[k])
    r.what = what
    if rdf is not None:
        r.rdf = rdf
    if sq is not None:
        r.sq = sq
    return r, (pos, it)


def request_footprint(r: Request) -> int:
    return lib().trj_request_footprint(C.byref(r))


class Server:
    """Continuous-batching request server (trj_srv_*)."""

    def __init__(self, slots: int = 4, budget_bytes: int = 0, device: int = 0):
        self._L = lib()
        h = C.c_void_p()
        rc = self._L.trj_srv_create(C.byref(h), device, slots, budget_bytes)
        if rc != TRJ_OK:
            raise TrjError(rc, "trj_srv_create failed")
        self._h = h
        self._keep = {}
        self._req = {}

    def close(self):
        if getattr(self, "_h", None):
            self._L.trj_srv_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != TRJ_OK:
            msg = self._L.trj_srv_last_error(self._h)
            raise TrjError(rc, f"{what}: {msg.decode() if msg else ''}")

    def submit(self, req: Request, keepalive=None) -> int:
        t = C.c_int64()
        self._check(self._L.trj_srv_submit(self._h, C.byref(req), C.byref(t)), "trj_srv_submit")
        self._keep[t.value] = (req, keepalive)
        self._req[t.value] = req
        return t.value

    def poll(self, ticket: int) -> bool:
        d = C.c_int32()
        self._check(self._L.trj_srv_poll(self._h, ticket, C.byref(d)), "trj_srv_poll")
        return bool(d.value)

    def result(self, ticket: int, lsmax: int = 0, nit: int = 0, nqmax: int = -1, nsp: int = 0):
        counts = np.zeros((nit, lsmax), dtype=np.int64) if lsmax > 0 else None
        sqf = sqfp = nq = None
        if nqmax >= 0:
            sqf = np.zeros(nqmax + 1)
            sqfp = np.zeros((nsp, nqmax + 1))
            nq = np.zeros(nqmax + 1, dtype=np.int64)
        self._check(self._L.trj_srv_result(self._h, ticket,
                                           _ptr(counts, C.c_int64) if counts is not None else None,
                                           _ptr(sqf, C.c_double) if sqf is not None else None,
                                           _ptr(sqfp, C.c_double) if sqfp is not None else None,
                                           _ptr(nq, C.c_int64) if nq is not None else None),
                    "trj_srv_result")
        self._keep.pop(ticket, None)
        self._req.pop(ticket, None)
        return counts, sqf, sqfp, nq

    def stats(self):
        q, r, b = C.c_int32(), C.c_int32(), C.c_int64()
        self._check(self._L.trj_srv_stats(self._h, C.byref(q), C.byref(r), C.byref(b)), "trj_srv_stats")
        return q.value, r.value, b.value

    def memory(self):
        """(bytes the slots hold now, the peak of that, requests that outgrew their reservation)."""
        a, b, m = C.c_int64(), C.c_int64(), C.c_int64()
        self._check(self._L.trj_srv_memory(self._h, C.byref(a), C.byref(b), C.byref(m)), "trj_srv_memory")
        return a.value, b.value, m.value
