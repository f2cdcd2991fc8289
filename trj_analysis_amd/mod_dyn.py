"""F(Q,t) tables and S(Q,w) of the reference's mod_dyn (src/modules/dynamics.cuf)
over the trj_hip C ABI: the GPU accumulates F(Q,t) and Fs(Q,t) over time
origins (Context.fqt_init, frames with TRJ_WANT_FQT); this module writes
fkt.dat, fskt.dat and sqw.dat as print_rtcor does (dynamics.cuf:246-330),
with the windowed cosine transform of src/core/fftwlib.f90 restated on
numpy's FFT (FFTW is not needed).

    fqt_init(ctx, qw, nbuffer, jump, ndyn, itlimit, box, ...)   dyn_init + the S(Q,w) q-set
    print_fqt(t, fkt, fskt, qw, itlimit, tmax, tunits, outdir)  print_rtcor (F(Q,t) part)
"""
from __future__ import annotations

import os
from typing import Optional

import numpy as np

from . import trjhip as T


def ndyn_of(nconf: int, nbuffer: int, jump: int) -> int:
    """dyn_init (dynamics.cuf:27-36): lags kept."""
    return nconf - (nbuffer - 1) * jump


def fftw1d(fin, dt: float, tmax: float):
    """fftwlib.f90:50-86 and 88-128: fin zero-padded to n = 2^(nu+1),
    nu = int(log(nin)/log(2) + 0.54), times the window 0.5 (1 - tanh(t - tmax))
    (tmax = n dt when |tmax| < 1e-6), first point halved (trapezoid), forward
    DFT times dt; bins 2..n/2 (1-based) are (F(k) + F(n-k+2)) dt / 2.  Returns
    (w, fout) for the first nin bins, w = (i-1) 2 pi / (n dt)."""
    fin = np.asarray(fin, dtype=np.float32)
    nin = len(fin)
    nu = int(np.float32(np.log(np.float32(nin)) / np.log(np.float32(2.0))) + np.float32(0.540))
    n = 2 ** (nu + 1)
    dt32 = np.float32(dt)
    dq = np.float32(2 * np.float32(np.pi)) / (np.float32(n) * dt32)
    x = np.zeros(n, dtype=np.complex128)
    x[:nin] = fin.astype(np.float64)
    tx = np.arange(n, dtype=np.float32) * dt32
    tm = np.float32(tmax)
    if abs(tm) < 1.0e-6:
        tm = np.float32(n) * dt32
    win = (np.float32(0.5) * (np.float32(1.0) - np.tanh(np.float32(1.0) * (tx - tm)))).astype(np.float32)
    x *= win.astype(np.float64)
    x[0] *= 0.5
    out = np.fft.fft(x)
    out[0] *= float(dt32)
    k = np.arange(1, n // 2)
    out[k] = (out[k] + out[n - k]) * float(dt32) / 2
    w = (np.arange(nin, dtype=np.float32) * dq).astype(np.float32)
    return w, out[:nin].real.astype(np.float32)


def _rows(values, width, prec, per):
    out = []
    for k in range(0, len(values), per):
        out.append("".join(f"{v:{width}.{prec}f}" for v in values[k:k + per]))
    return "\n".join(out)


def print_fqt(t, fkt, fskt, qw, itlimit: int, tmax: float, tunits: str = "lj", outdir: str = ".") -> list:
    """print_rtcor's F(Q,t) part (dynamics.cuf:246-330): fkt.dat and fskt.dat
    for t(i), i <= itlimit, and sqw.dat with S(Q,w) and Ss(Q,w) of fftw1d.
    t: times of the ndyn lags (t - t(1) is written); fkt, fskt (nqw, ndyn)
    normalised (Context.fqt_results)."""
    t = np.asarray(t, dtype=np.float64)
    t = t - t[0]
    fkt = np.asarray(fkt)
    fskt = np.asarray(fskt)
    nqw, ndyn = fkt.shape
    lj = tunits == "lj"
    tl = "t(tau)" if lj else "t(ps)"
    qs = "".join(f"F({q:6.4f},t)    " for q in qw).rstrip()
    qss = "".join(f"Fs({q:6.4f},t)    " for q in qw).rstrip()
    p1, p2, p3 = (os.path.join(outdir, f) for f in ("fkt.dat", "fskt.dat", "sqw.dat"))
    dt = t[1] - t[0] if ndyn > 1 else 1.0
    sqw = np.zeros((ndyn, nqw), dtype=np.float32)
    ssqw = np.zeros((ndyn, nqw), dtype=np.float32)
    w = None
    for i in range(nqw):
        w, sqw[:, i] = fftw1d(fkt[i], dt, tmax)
        _, ssqw[:, i] = fftw1d(fskt[i], dt, tmax)
    fconv = np.float32(4.135)
    with open(p1, "w") as f1, open(p2, "w") as f2, open(p3, "w") as f3:
        f1.write(f"#  {tl}         {qs}\n")
        f2.write(f"#  {tl}         {qss}\n")
        sw = "".join(f"S({q:6.4f},w)    Ss({q:6.4f},w)    " for q in qw).rstrip()
        f3.write(("#  w(1/tau)     " if lj else "#  w(THz)       w(meV)     ") + sw + "\n")
        for i in range(ndyn):
            if i + 1 <= itlimit:
                f1.write(f"{t[i]:10.3f}" + _rows(list(fkt[:, i]), 15, 5, 8) + "\n")
                f2.write(f"{t[i]:10.3f}" + _rows(list(fskt[:, i]), 15, 5, 8) + "\n")
            vals = [v for j in range(nqw) for v in (sqw[i, j], ssqw[i, j])]
            wi = float(w[i]) / (2 * np.pi)
            if lj:
                f3.write(f"{wi:10.3f}" + _rows(vals, 15, 5, 8) + "\n")
            else:
                f3.write(f"{wi:10.3f}{float(fconv) * wi:10.3f}" + _rows(vals, 15, 5, 8) + "\n")
    return [p1, p2, p3]
