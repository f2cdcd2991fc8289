"""The reference's mod_sq interface over the trj_hip C ABI.

Same procedure names and argument meaning as src/modules/sq.cuf and the S(Q)
writer of src/core/log.f90:

  sq_init(Nsites, nsp, ...)          sq.cuf:150-369
  SQcalc(r, itype, sidel)            sq.cuf:58-126 (one frame; accumulates)
  sq_transfer_gpu_cpu()              sq.cuf:128-148 (returns sqf, sqfp, nq)
  printSQ(Nsites, ...)               log.f90:317-412 (sq.dat, sqmix.dat)
  sq_clear()                         sq.cuf:774

With flags=TRJ_SQ_CHARGE (the reference's ex_qc) the charge channel S_qq is
summed too (sq.cuf:424-432, 563-582, 143-145) and printSQ writes its column
after S_NN (log.f90:351-360, 385-393).

Values of the reference's FP32 bookkeeping (dq, nqmax, nqmin, lmax, qmin after
the lattice adjust) are returned by sq_init as the reference computes them.
"""
from __future__ import annotations

import os
from typing import Optional

import numpy as np

from . import trjhip as T
from .mod_rdf import context

_state: dict = {}


def sq_init(Nsites: int, nsp: int, *, ndim: int, qmax: float, qmin: Optional[float], bsc, sidel,
            qset: int = T.TRJ_QSET_REF, grid=(0, 0, 0), precision: str = "fp64", flags: int = 0,
            ctx: Optional[T.Context] = None) -> T.SqInfo:
    """sq_init (sq.cuf:150-369) on the first frame's box; returns the q-grid info
    (dq, nqmax, nqmin, lmax, qmin_eff)."""
    c = ctx or context()
    info = c.sq_init(T.sq_params(ndim, nsp, qmax, qmin, bsc, qset, grid, precision, flags), sidel)
    _state.update(ctx=c, info=info, nsp=nsp, Nsites=Nsites, nframes=0,
                  charges=bool(flags & T.TRJ_SQ_CHARGE))
    return info


def SQcalc(r: np.ndarray, itype: np.ndarray, sidel, qcharge: Optional[np.ndarray] = None):
    """SQcalc (sq.cuf:58-126) for one frame: r (Nsites, ndim), itype 1..nsp,
    qcharge the atoms' charges (required with TRJ_SQ_CHARGE)."""
    c = _state["ctx"]
    r = np.asarray(r)
    if r.shape[0] != len(itype):
        r = r.T
    c.frame(np.ascontiguousarray(r), itype, sidel, what=T.TRJ_WANT_SQ, charge=qcharge)
    _state["nframes"] += 1


def sq_transfer_gpu_cpu():
    """sq_transfer_gpu_cpu (sq.cuf:128-148): (sqf[0:nqmax], sqfp[nsp][0:nqmax], nq)."""
    return _state["ctx"].sq_shells()


def sq_clear():
    """sq_clear (sq.cuf:774): drop the accumulators."""
    c = _state.get("ctx")
    if c is not None:
        c.reset()
    _state.clear()


def _fmt15(values) -> str:
    """Fortran '(15f16.7)' with format reversion."""
    out = []
    for k in range(0, len(values), 15):
        out.append("".join(f"{v:16.7f}" for v in values[k:k + 15]))
    return "\n".join(out)


def _printSQ_slices(c, nconf: int, outdir: str) -> list:
    """log.f90:333-345, 366-378: sqxy.dat, sqpxy_j-j.dat and sqxy_qq.dat."""
    from .mod_rdf import _slice_head
    nsp = _state["nsp"]
    qq = _state.get("charges", False)
    cols, pcols, qcols = c.sq_xy_columns(nconf, charges=qq)
    head = _slice_head("#     Q        ", "S_xy(Q,", c.zslice, 3)
    paths = []
    path = os.path.join(outdir, "sqxy.dat")
    with open(path, "w") as f:
        f.write(head + "\n")
        for row in cols:
            f.write(_fmt15(row) + "\n")
    paths.append(path)
    if qq:
        path = os.path.join(outdir, "sqxy_qq.dat")
        with open(path, "w") as f:
            f.write(_slice_head("#     Q        ", "S_xy_qq(Q,", c.zslice, 3) + "\n")
            for row in qcols:
                f.write(_fmt15(row) + "\n")
        paths.append(path)
    for j in range(nsp):
        path = os.path.join(outdir, f"sqpxy_{j + 1}-{j + 1}.dat")
        with open(path, "w") as f:
            f.write(head + "\n")
            for row in pcols[j]:
                f.write(_fmt15(row) + "\n")
        paths.append(path)
    return paths


def printSQ(Nsites: int, nconf: int = 0, outdir: str = ".") -> list:
    """printSQ (log.f90:317-412): writes sq.dat and sqmix.dat (sqxy*.dat with
    z-slices); returns their paths."""
    c = _state["ctx"]
    nsp = _state["nsp"]
    if c.nslice > 0:
        return _printSQ_slices(c, nconf, outdir)
    cols, mix = c.sq_columns(nconf)
    if _state.get("charges", False):   # S_qq after S_NN (log.f90:351-360, 385-393)
        cols = np.insert(cols, 2, c.sq_qq_column(nconf), axis=1)
        qh = "          S_qq(Q)"
    else:
        qh = ""
    p_sq = os.path.join(outdir, "sq.dat")
    p_mix = os.path.join(outdir, "sqmix.dat")
    nv = cols.shape[1] - 1
    with open(p_sq, "w") as f:
        if nv >= 6:
            f.write("#           Q        S_NN(Q)" + qh + "          S_cc(Q)       S_11(Q)        S_22(Q)"
                    "           S_12(Q)         n(Q)\n")
        else:
            f.write("#           Q       S_NN(Q)" + qh + "          n(Q)\n")
        for row in cols:
            f.write("".join(f"{v:15.7f}" for v in row[:nv]) + f"{int(row[nv]):12d}\n")
    with open(p_mix, "w") as f:
        f.write(("#       Q" + " " * 14 + "".join(f"S_{j}{j}(Q)" + " " * 9 for j in range(1, nsp + 1))).rstrip() + "\n")
        for row in mix:
            f.write("".join(f"{v:16.7f}" for v in row) + "\n")
    return [p_sq, p_mix]


def printSQcl(nconf: int, outdir: str = ".", ctx: Optional[T.Context] = None) -> str:
    """sqcl.dat (log.f90:637-642) of a context initialised with TRJ_SQ_CENTRES
    and fed the cluster centres (sqfnDcl): i dq, sqfcl(i)/(Nconf nq(i)) for
    i = 1..nqmin."""
    c = ctx or _state["ctx"]
    if nconf <= 0:
        raise ValueError("printSQcl needs Nconf (the number of frames)")
    sqf, _, nq = c.sq_shells()
    nc = nconf
    info = c.sq_info
    path = os.path.join(outdir, "sqcl.dat")
    with open(path, "w") as f:
        f.write("#      Q            S_cl-cl(Q)\n")
        for i in range(1, info.nqmin + 1):
            q = float(np.float32(i) * np.float32(info.dq))
            v = sqf[i] / (nc * nq[i]) if nq[i] else 0.0
            f.write(f"{q:15.7f}{v:15.7f}\n")
    return path

