"""F(Q,t) tables of the reference's mod_dyn (src/modules/dynamics.cuf) over the
trj_hip C ABI: the GPU accumulates F(Q,t) and Fs(Q,t) over time origins
(Context.fqt_init, frames with TRJ_WANT_FQT); this module writes fkt.dat and
fskt.dat as print_rtcor does (dynamics.cuf:246-330).  The S(Q,w) table of
print_rtcor (sqw.dat, through the FFTW transform of src/core/fftwlib.f90)
is not written: the transform belongs to the dynamics module, outside this
library's scope.

    ndyn_of(nconf, nbuffer, jump)                          dyn_init's lag count
    print_fqt(t, fkt, fskt, qw, itlimit, tunits, outdir)   print_rtcor (F(Q,t) part)
"""
from __future__ import annotations

import os

import numpy as np


def ndyn_of(nconf: int, nbuffer: int, jump: int) -> int:
    """dyn_init (dynamics.cuf:96): lags kept, nconf - (nbuffer - 1) * jump."""
    return nconf - (nbuffer - 1) * jump


def _rows(values, width, prec, per):
    out = []
    for k in range(0, len(values), per):
        out.append("".join(f"{v:{width}.{prec}f}" for v in values[k:k + per]))
    return "\n".join(out)


def print_fqt(t, fkt, fskt, qw, itlimit: int, tunits: str = "lj", outdir: str = ".") -> list:
    """print_rtcor's F(Q,t) part (dynamics.cuf:246-330): fkt.dat and fskt.dat
    for t(i), i <= itlimit.  t: times of the ndyn lags (t - t(1) is written);
    fkt, fskt (nqw, ndyn) normalised (Context.fqt_results)."""
    t = np.asarray(t, dtype=np.float64)
    t = t - t[0]
    fkt = np.asarray(fkt)
    fskt = np.asarray(fskt)
    nqw, ndyn = fkt.shape
    tl = "t(tau)" if tunits == "lj" else "t(ps)"
    qs = "".join(f"F({q:6.4f},t)    " for q in qw).rstrip()
    qss = "".join(f"Fs({q:6.4f},t)    " for q in qw).rstrip()
    p1, p2 = (os.path.join(outdir, f) for f in ("fkt.dat", "fskt.dat"))
    with open(p1, "w") as f1, open(p2, "w") as f2:
        f1.write(f"#  {tl}         {qs}\n")
        f2.write(f"#  {tl}         {qss}\n")
        for i in range(min(ndyn, itlimit)):
            f1.write(f"{t[i]:10.3f}" + _rows(list(fkt[:, i]), 15, 5, 8) + "\n")
            f2.write(f"{t[i]:10.3f}" + _rows(list(fskt[:, i]), 15, 5, 8) + "\n")
    return [p1, p2]
