"""The reference's mod_rdf interface over the trj_hip C ABI.

Same procedure names and argument meaning as src/modules/rdf.cuf and the RDF
writer of src/core/log.f90, with the module globals (mod_common, mod_input)
replaced by explicit arguments and one module-level context:

  RDF_init(nsp, ...)                 rdf.cuf:51-122
  RDFcomp(Nsites, Iconf, r, itype, sidel)   rdf.cuf:184-353
  rdf_clear()                        rdf.cuf:124-182
  printrdf(rcl, lsmax, ...)          log.f90:414-596 (gmixsim.dat / gmixsim<k>.dat)

With charges (the reference's ex_qc: the trajectory holds a q variable),
RDF_init(..., charges=True) adds the charge-weighted histogram and printrdf
its g_qq column (rdf.cuf:466-472, 270-274; log.f90:466-468, 501-503, 565-575).

With z-slices set on the context (the reference's twoDstruc_3D,
Context.set_slices before RDF_init), the histograms are rdfxy's and printrdf
writes gxy_j-l.dat (and gxy_cc.dat with charges) instead (log.f90:424-437,
494-561).

Errors raise trjhip.TrjError (the reference prints and stops, rdf.cuf:241-242).
Histograms are exact int64 counts (the reference accumulates them in FP32,
rdf.cuf:306, common.cuf:90).
"""
from __future__ import annotations

import os
from typing import Optional

import numpy as np

from . import trjhip as T

_ctx: Optional[T.Context] = None
_state: dict = {}


def context() -> T.Context:
    global _ctx
    if _ctx is None:
        _ctx = T.Context(int(os.environ.get("TRJ_DEVICE", "0")))
    return _ctx


def RDF_init(nsp: int, *, ndim: int, deltar: float, rcrdf: float, sidel, periodic=(True, True, True),
             charges: bool = False, nrandom: int = 0, ctx: Optional[T.Context] = None) -> int:
    """RDF_init (rdf.cuf:51-122): returns lsmax; sidel is the first frame's box.
    charges=True also accumulates g_qq (frames then need qcharge); nrandom > 0
    samples the number fluctuations around that many fixed points (s2n.dat)."""
    c = ctx or context()
    flags = T.TRJ_RDF_CHARGE if charges else 0
    st = c.rdf_init(T.rdf_params(ndim, nsp, deltar, rcrdf, periodic, flags=flags), sidel)
    if nrandom > 0:
        c.rdf_sampling(nrandom)
    _state.update(ctx=c, lsmax=st.lsmax, nit=st.nit, nsp=nsp, ndim=ndim, deltar=deltar,
                  rmingr=None, nframes=0, charges=charges, nrandom=nrandom)
    return st.lsmax


def RDFcomp(Nsites: int, Iconf: int, r: np.ndarray, itype: np.ndarray, sidel, units: str = "lj",
            qcharge: Optional[np.ndarray] = None):
    """RDFcomp (rdf.cuf:184-353) for one frame: r is (Nsites, ndim) (or the
    reference's r(ndim,N) transposed), itype 1..nsp, qcharge the atoms'
    charges (required after RDF_init(..., charges=True)).  On the first frame
    the suggested rcl of rdf.cuf:344-352 is computed and kept in rdf_rmingr()."""
    c = _state["ctx"]
    r = np.asarray(r)
    if r.shape[0] != Nsites:
        r = r.T
    c.frame(np.ascontiguousarray(r), itype, sidel, what=T.TRJ_WANT_RDF, charge=qcharge)
    _state["nframes"] += 1
    if Iconf == 1 and c.nslice == 0:  # rdf.cuf:343 (not with slices)
        _state["rmingr"] = c.rdf_first_minimum()


def rdf_rmingr() -> Optional[float]:
    return _state.get("rmingr")


def histomix() -> np.ndarray:
    """The accumulated histogram, histomix(lsmax, nsp, nsp) as (nsp, nsp, lsmax), symmetric."""
    c = _state["ctx"]
    counts = c.rdf_counts()
    nsp, ls = _state["nsp"], _state["lsmax"]
    h = np.zeros((nsp, nsp, ls), dtype=np.int64)
    k = 0
    for j in range(nsp):
        for l in range(j, nsp):
            h[j, l] = counts[k]
            h[l, j] = counts[k]
            k += 1
    return h


def rdf_clear():
    """rdf_clear (rdf.cuf:124-182): drop the accumulators."""
    c = _state.get("ctx")
    if c is not None:
        c.reset()
    _state.clear()


def _fmt_records(values, width, prec, per_record=16):
    """Fortran '(16fW.D)' with format reversion: 16 values per line."""
    out = []
    for k in range(0, len(values), per_record):
        out.append("".join(f"{v:{width}.{prec}f}" for v in values[k:k + per_record]))
    return "\n".join(out)


def _ri(i: int, dr: float) -> float:
    """ri = i*deltar in FP32, as log.f90:485."""
    return float(np.float32(i) * np.float32(dr))


def _slice_head(first: str, label: str, zslice, gap: int) -> str:
    """'(a,3x,16(label,f7.2,')',gap x:))' without trailing blanks."""
    return (first + "   " + "".join(f"{label}{z:7.2f})" + " " * gap for z in zslice)).rstrip()


def _printrdf_slices(c, nconf: int, outdir: str) -> list:
    """log.f90:424-437, 494-561: gxy_j-l.dat per pair j <= l and gxy_cc.dat."""
    nsp = _state["nsp"]
    g = c.rdf_xy_gmix(nconf)                        # (pair, lsmax-2, nslice)
    dr = _state["deltar"]
    head = _slice_head("#     r        ", "g_xy(r,", c.zslice, 3)
    paths = []
    pair = 0
    for j in range(1, nsp + 1):
        for l in range(j, nsp + 1):
            path = os.path.join(outdir, f"gxy_{j}-{l}.dat")
            with open(path, "w") as f:
                f.write(head + "\n")
                for i in range(g.shape[1]):
                    f.write(_fmt_records([_ri(i + 1, dr)] + list(g[pair, i]), 16, 7, 15) + "\n")
            paths.append(path)
            pair += 1
    if _state.get("charges", False):
        gc = c.rdf_xy_gcc(nconf)
        path = os.path.join(outdir, "gxy_cc.dat")
        with open(path, "w") as f:
            f.write(_slice_head("#     r        ", "g_xy_cc(r,", c.zslice, 4) + "\n")
            for i in range(gc.shape[0]):
                f.write(_fmt_records([_ri(i + 1, dr)] + list(gc[i]), 16, 7, 15) + "\n")
        paths.append(path)
    return paths


def printrdf(rcl: float = 0.0, lsmax: Optional[int] = None, nconf: int = 0, outdir: str = ".") -> list:
    """printrdf (log.f90:414-596): writes gmixsim.dat (nsp <= 6) or gmixsim<k>.dat
    and returns the written paths.  nconf <= 0 uses the number of frames processed."""
    c = _state["ctx"]
    if c.nslice > 0:
        return _printrdf_slices(c, nconf, outdir)
    paths = []
    if _state.get("nrandom", 0) > 0:  # log.f90:449, 482, 526-533
        s2 = c.rdf_s2n(nconf)
        path = os.path.join(outdir, "s2n.dat")
        with open(path, "w") as f:
            f.write("#       r" + " " * 16 + "s2n\n")
            for i, v in enumerate(s2):
                f.write(_fmt_records([_ri(i + 1, _state["deltar"]), v], 16, 5, 18) + "\n")
        paths.append(path)
    nsp, nit = _state["nsp"], _state["nit"]
    g = c.rdf_gmix(nconf)
    qq = _state.get("charges", False)
    gqq = c.rdf_gqq(nconf) if qq else None
    if qq and len(gqq) != len(g):
        raise T.TrjError(T.TRJ_ESTATE, "g_qq and g_ab row counts differ")
    qhead = "g_qq(r)" + " " * 9 if qq else ""
    if nsp <= 6:
        pairs = [(j, k) for j in range(1, nsp + 1) for k in range(j, nsp + 1)]
        head = ("#       r        " + " " * 9 + qhead + "".join(f"g_{j}{k}(r)" + " " * 9 for j, k in pairs)).rstrip()
        path = os.path.join(outdir, "gmixsim.dat")
        with open(path, "w") as f:
            f.write(head + "\n")
            for i, row in enumerate(g):
                if qq:   # log.f90:565-566: '(16f16.6)' ri, gchch, gmix
                    f.write(_fmt_records([row[0], gqq[i]] + list(row[1:]), 16, 6) + "\n")
                else:
                    f.write(_fmt_records(list(row), 16, 5) + "\n")
        paths.append(path)
    else:
        full = np.zeros((len(g), nsp, nsp))
        col = 1
        for j in range(nsp):
            for k in range(j, nsp):
                full[:, j, k] = g[:, col]
                full[:, k, j] = g[:, col]
                col += 1
        for k in range(1, nsp + 1):
            lead = qhead if k == 1 else ""   # log.f90:474-476, 572-573: g_qq in the first file
            head = ("#       r        " + " " * 9 + lead
                    + "".join(f"g_{k}{j}(r)" + " " * 9 for j in range(1, nsp + 1))).rstrip()
            path = os.path.join(outdir, f"gmixsim{k}.dat")
            with open(path, "w") as f:
                f.write(head + "\n")
                for i, row in enumerate(g):
                    extra = [gqq[i]] if (qq and k == 1) else []
                    f.write(_fmt_records([row[0]] + extra + list(full[i, k - 1]), 16, 5) + "\n")
            paths.append(path)
    return paths
