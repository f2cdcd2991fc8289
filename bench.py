#!/usr/bin/env python3
"""Benchmark of the structural hot path (RDF + S(Q)) on one MI355X.

    python bench.py [--steps K] [--warmup W] [--workload lattice_262k]

A step is one frame through the whole path: ingest (wrap, species sort), the
cell-list pair histogram g_ab(r) and the S(Q) contraction with its shell sums.
Frames are generated from a seed and copied to HBM before the timed region;
the timed region runs K frames back to back on the context stream and ends
with a device synchronisation.  One JSON line is printed.

`value` is frames/s of the lattice_262k workload (262144 atoms, FP64, the
reference q-set to qmax = 20).  Fields besides the driver's contract:
  pairs_per_s         in-range pairs binned per second (RDF)
  q_atom_terms_per_s  (q-vector, atom) phase terms per second (S(Q))
  latency_ms_p50/p99  per-frame latency, frames submitted one at a time
  kernel_ms           mean per-frame time of each stage (HIP events on the
                      context stream)
  roofline            the dominant kernel (the S(Q) contraction) against the
                      FP64 peak: algorithmic flops 8*N*Nq per launch / mean
                      launch time
  cpu_baseline        the CPU port on the host cores, on a bounded sample of
                      the same workload, extrapolated to frames/s
  north_star          the 1M-atom, 3-species frame (rc = 5, 64^3 half-space
                      q-grid): the plain FP64 per-frame path against FP16x3
                      (serial and overlapped), the 16-bit contraction's
                      roofline, FP16x3's per-shell error against FP64
  configs0            the reference's CPU case (10 frames, N = 4000, two
                      species) in full on the host cores and on the GPU
  batch               uneven frames in graph-replayed batched launches
                      against the one-frame-per-launch path
  serving             the request server against the per-request path on
                      Poisson arrivals (p50/p99 latency, throughput)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "frames/s + pairs/s (RDF), q·atom terms/s (S(Q)); p50/p99 frame latency"
PEAK = {"fp64": 78.6e12, "fp32": 157.3e12, "fp16x3": 2.5e15 / 3, "bf16x3": 2.5e15 / 3}
# Nominal dense rate of the MFMA instruction each path issues (PEAK divides the
# 16-bit one by the three products of a split-precision real product).
NOMINAL_MFMA = {"fp64": 78.6e12, "fp16x3": 2.5e15, "bf16x3": 2.5e15}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_frames(workload: str, count: int):
    from trj_analysis_amd import workloads as W
    if workload == "lattice_262k":
        wl = W.LATTICE_262K
        frames = list(W.lattice_262k(nframes=count))
    elif workload == "lj_binary_4k":
        wl = W.LJ_BINARY_4K
        frames = list(W.lj_binary_4k(nframes=count))
    elif workload == "uniform_1m":
        wl = W.UNIFORM_1M
        frames = [W.uniform_1m(seed=2 + k) for k in range(count)]
    else:
        raise SystemExit(f"unknown workload {workload}")
    return wl, frames


def cpu_baseline(wl, frame, n_in_range_pairs: float, nq_terms: int, budget_s: float):
    """Oracle (the CPU port of the reference algorithm) on the host cores."""
    from oracle import oracle as O
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    box, per = frame.box, (1, 1, 1)
    # RDF: one full frame through the OpenMP linked-cell oracle.
    t0 = time.perf_counter()
    counts = O.rdf_counts(frame.pos, frame.itype, 3, wl.nsp, box, per, wl.rcrdf, wl.deltar)
    t_rdf = time.perf_counter() - t0
    # S(Q): FP64 direct sums (double sincos, OpenMP over vectors) over a sample
    # of the q-set.
    pos = O.wrap(frame.pos, box, per)
    rng = np.random.default_rng(0)
    kv = rng.integers(0, 8, size=(64, 3)).astype(np.int32)
    t0 = time.perf_counter()
    O.sq_rho_f64(pos, frame.itype, wl.nsp, box, kv)
    t64 = time.perf_counter() - t0
    m = int(max(64, min(200000, 64 * max(0.0, budget_s - t_rdf) / max(t64, 1e-6))))
    kv = rng.integers(0, 32, size=(m, 3)).astype(np.int32)
    t0 = time.perf_counter()
    O.sq_rho_f64(pos, frame.itype, wl.nsp, box, kv)
    t_m = time.perf_counter() - t0
    n = len(frame.itype)
    terms_rate = m * n / t_m
    t_sq = nq_terms / terms_rate
    pairs_rate = float(counts.sum()) / t_rdf
    fps = 1.0 / (t_rdf + t_sq)
    return {
        "value": fps, "unit": "frames/s", "cores": cores, "kind": "port",
        "sample": (f"one {n}-atom frame: RDF in full (linked cells, OpenMP, {t_rdf:.2f} s); S(Q) "
                   f"FP64 direct sums over {m} q-vectors ({t_m:.2f} s), extrapolated to the "
                   f"{nq_terms // n} vectors of a frame"),
        "pairs_per_s": pairs_rate, "q_atom_terms_per_s": terms_rate,
    }


def bench_batch(nf: int) -> dict:
    """uneven_batch: nf frames of N log-uniform in [1e4, 1e6], 2-4 species,
    float32, rcrdf=4, deltar=0.01; per-frame pair histograms.  Timed: the
    batched graph-replayed path (autotuned on the first frames) against the
    one-frame-per-launch path on the same frames; results compared bit for bit."""
    import torch
    from trj_analysis_amd import trjhip as T
    from trj_analysis_amd import workloads as W

    log(f"generating {nf} uneven frames")
    frames = [W.uneven_frame(1000 + k) for k in range(nf)]
    natoms = sum(len(f.itype) for f in frames)
    pos = torch.from_numpy(np.concatenate([f.pos for f in frames]).astype(np.float32)).cuda()
    it = torch.from_numpy(np.concatenate([f.itype for f in frames]).astype(np.int32)).cuda()
    offs = np.concatenate([[0], np.cumsum([len(f.itype) for f in frames])])
    torch.cuda.synchronize()
    wl = W.UNEVEN_BATCH
    fa = T.batch_frames((len(f.itype), f.nsp, f.box, f.periodic) for f in frames)

    # one frame per launch (the per-frame GPU path), each frame its own request
    ctx = T.Context(0)
    plain = []

    def per_frame(collect):
        for k, f in enumerate(frames):
            ctx.rdf_init(T.rdf_params(3, f.nsp, wl.deltar, wl.rcrdf, f.periodic, flags=T.TRJ_RDF_NO_AUTOTUNE),
                         f.box)
            ctx.frame_device(pos.data_ptr() + 12 * int(offs[k]), T.TRJ_F32, T.TRJ_AOS,
                             it.data_ptr() + 4 * int(offs[k]), len(f.itype), f.box, T.TRJ_WANT_RDF)
            c = ctx.rdf_counts()
            if collect:
                plain.append(c)
    per_frame(False)  # warm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    per_frame(True)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    ctx.close()

    b = T.Batch(T.batch_params(3, wl.deltar, wl.rcrdf, use_graph=True))
    chosen = b.autotune(fa, pos.data_ptr(), T.TRJ_F32, it.data_ptr())
    log(f"autotuned: frames_per_launch={chosen.frames_per_launch} workgroups={chosen.workgroups} "
        f"waves_per_cell={chosen.waves_per_cell}")
    b.run(fa, pos.data_ptr(), T.TRJ_F32, it.data_ptr(), want_hist=False)  # warm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    got, gpu_ms = b.run(fa, pos.data_ptr(), T.TRJ_F32, it.data_ptr())
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    exact = all(np.array_equal(got[k], plain[k]) for k in range(nf))
    pairs = sum(int(c.sum()) for c in plain)
    nog = T.Batch(T.batch_params(3, wl.deltar, wl.rcrdf, chosen.frames_per_launch, chosen.waves_per_cell,
                                 chosen.workgroups, use_graph=False))
    nog.run(fa, pos.data_ptr(), T.TRJ_F32, it.data_ptr(), want_hist=False)
    t0 = time.perf_counter()
    nog.run(fa, pos.data_ptr(), T.TRJ_F32, it.data_ptr(), want_hist=False)
    t_nograph = time.perf_counter() - t0
    b.close()
    nog.close()
    sq = {p: bench_batch_sq(frames, pos, it, offs, p) for p in ("fp16x3", "fp64")}
    return {
        "frames": nf, "atoms": natoms, "rcrdf": wl.rcrdf, "deltar": wl.deltar,
        "inputs": "HBM-resident float32",
        "autotuned": {"frames_per_launch": chosen.frames_per_launch, "workgroups": chosen.workgroups,
                      "waves_per_cell": chosen.waves_per_cell},
        "frames_per_s": nf / t_batch, "ms_per_frame": t_batch * 1e3 / nf,
        "pairs_per_s": pairs / t_batch,
        "plain_frames_per_s": nf / t_plain,
        "speedup_vs_plain": t_plain / t_batch,
        "no_graph_frames_per_s": nf / t_nograph,
        "batch_gpu_ms": gpu_ms,
        "bit_exact_vs_plain": exact,
        "sq": sq,
    }


def bench_batch_sq(frames, pos, it, offs, precision: str) -> dict:
    """S(Q) of the uneven frames on the 64^3 half-space grid of each frame's own
    box: the batched path (trj_batch_sq_run, graph-replayed groups) against
    the per-frame path (trj_sq_init on the frame's box, untimed, then one
    trj_frame_device and trj_sq_shells, timed) on the same HBM-resident
    frames; the largest relative difference of any shell sum between them."""
    import torch
    from trj_analysis_amd import trjhip as T

    grid = (64, 64, 64)
    prec = T.PRECISIONS[precision]
    nf = len(frames)
    bsc = [[1.0, 0.6, -0.3, 0.4][:f.nsp] for f in frames]
    ctx = T.Context(0)
    plain = []
    t_plain = 0.0
    for rep in range(2):  # the first pass warms up
        t_plain = 0.0
        plain = []
        for k, f in enumerate(frames):
            ctx.sq_init(T.sq_params(3, f.nsp, 1e9, bsc=bsc[k], qset=T.TRJ_QSET_HALFSPACE, grid=grid,
                                    precision=precision, flags=T.TRJ_SQ_NO_AUTOTUNE), f.box)
            ctx.sync()
            t0 = time.perf_counter()
            ctx.frame_device(pos.data_ptr() + 12 * int(offs[k]), T.TRJ_F32, T.TRJ_AOS,
                             it.data_ptr() + 4 * int(offs[k]), len(f.itype), f.box, T.TRJ_WANT_SQ)
            plain.append(ctx.sq_shells())
            t_plain += time.perf_counter() - t0
    ctx.close()
    fa = T.batch_frames((len(f.itype), f.nsp, f.box, (1, 1, 1)) for f in frames)
    b = T.Batch(T.batch_params(3, 0.01, 4.0, use_graph=True))
    q = T.batch_sq_params(grid, prec, 1e9)
    b.sq_run(q, fa, pos.data_ptr(), T.TRJ_F32, it.data_ptr(), bsc=bsc)  # warm (captures)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    got, gpu_ms = b.sq_run(q, fa, pos.data_ptr(), T.TRJ_F32, it.data_ptr(), bsc=bsc)
    t_batch = time.perf_counter() - t0
    b.close()
    rel = 0.0
    nq_equal = True
    for (sqf, sqfp, nq), (wf, wfp, wnq) in zip(got, plain):
        nq_equal = nq_equal and np.array_equal(nq, wnq)
        m = wnq > 0
        rel = max(rel, float(np.max(np.abs(sqf[m] - wf[m]) / np.abs(wf[m]))),
                  float(np.max(np.abs(sqfp[:, m] - wfp[:, m]) / np.abs(wfp[:, m]))))
    natoms = int(offs[-1])
    terms = natoms * (grid[0] * grid[1] * grid[2] // 2)
    log(f"batched S(Q) {precision}: {nf / t_batch:.1f} frames/s vs per-frame {nf / t_plain:.1f}; "
        f"max shell rel diff {rel:.2e}")
    return {
        "precision": precision, "grid": list(grid), "qset": "halfspace",
        "frames_per_s": nf / t_batch, "ms_per_frame": t_batch * 1e3 / nf, "batch_gpu_ms": gpu_ms,
        "q_atom_terms_per_s": terms / t_batch,
        "plain_frames_per_s": nf / t_plain, "speedup_vs_plain": t_plain / t_batch,
        "nq_equal": nq_equal, "max_shell_rel_diff_vs_plain": rel,
        "note": "plain: per frame trj_sq_init on its box (untimed, default K-split: no first-frame "
                "autotuning per frame), then trj_frame_device + trj_sq_shells (timed); batch: "
                "trj_batch_sq_run over all frames (graph-replayed groups of 64)",
    }


def bench_serving(rates, nreq: int, slots: int) -> dict:
    """serving: Poisson streams of RDF + S(Q) requests (bulk or z-slab frames, N
    log-uniform in [1e4, 2e6], 1-4 species, float32, 64^3 half-space q-grid,
    FP16x3 contraction) at the offered rates; the continuous-batching server
    against the plain per-request path on the same arrivals."""
    from trj_analysis_amd import serve as S
    out_rates = {}
    for rate in rates:
        jobs = S.make_jobs(nreq, rate=rate, seed=11, precision="fp16x3")
        log(f"rate {rate}/s: {len(jobs)} requests, mean N {np.mean([len(j.frame.itype) for j in jobs]):.0f}")
        S.serve_stream(jobs[:4], slots=slots)  # warm
        lat_s, res_s, span_s = S.serve_stream(jobs, slots=slots)
        lat_q, res_q, span_q = S.serve_sequential(jobs)
        exact = all(np.array_equal(a[0], b[0]) for a, b in zip(res_s, res_q))
        rel = max(float(np.max(np.abs(a[1] - b[1]) / np.maximum(np.abs(b[1]), 1e-300)))
                  for a, b in zip(res_s, res_q))
        out_rates[str(rate)] = {
            "server": {"throughput_req_s": len(jobs) / span_s, "p50_ms": 1e3 * float(np.percentile(lat_s, 50)),
                       "p99_ms": 1e3 * float(np.percentile(lat_s, 99))},
            "per_request": {"throughput_req_s": len(jobs) / span_q,
                            "p50_ms": 1e3 * float(np.percentile(lat_q, 50)),
                            "p99_ms": 1e3 * float(np.percentile(lat_q, 99))},
            "counts_bit_exact": exact, "sq_max_rel_diff": rel,
        }
        log(json.dumps(out_rates[str(rate)]))
    return {"requests_per_rate": nreq, "slots": slots, "rates": out_rates, "rcrdf": S.RCRDF,
            "deltar": S.DELTAR, "q_grid": list(S.GRID), "precision": "fp16x3"}


# FP32 VALU lane operations per second of the MI355X (157.3 TFLOP/s counts an
# FMA as two flops; one lane operation per lane per instruction).
VALU_PEAK = 157.3e12 / 2
# The FP32 work one candidate pair needs at least: three coordinate
# differences, r^2 in three multiply-adds, one comparison with the cutoff.
RDF_OPS_PER_CANDIDATE = 7


def rdf_candidates(pos, box, rc) -> float:
    """Candidate pairs of the pair kernel's cell walk on a periodic 3D frame:
    cells of side >= rc (floor(L / rc) per axis, at most 160), each cell's own
    pairs j > i plus every atom against the 13 forward neighbour cells."""
    nc = [max(1, min(160, int(np.floor(L / (rc * (1.0 + 1e-6)))))) for L in box]
    idx = [np.minimum((np.mod(pos[:, k], box[k]) * (nc[k] / box[k])).astype(np.int64), nc[k] - 1)
           for k in range(3)]
    flat = (idx[2] * nc[1] + idx[1]) * nc[0] + idx[0]
    cnt = np.bincount(flat, minlength=nc[0] * nc[1] * nc[2]).reshape(nc[2], nc[1], nc[0]).astype(np.float64)
    c = float((cnt * (cnt - 1.0) / 2.0).sum())
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                if dz > 0 or (dz == 0 and dy > 0) or (dz == 0 and dy == 0 and dx > 0):
                    c += float((cnt * np.roll(cnt, (-dz, -dy, -dx), axis=(0, 1, 2))).sum())
    return c


def rdf_valu_roofline(frame, rc, rdf_ms) -> dict:
    """The RDF stage (cell build, pair kernel, merge) priced as candidate
    lane operations against the FP32 VALU peak."""
    cand = rdf_candidates(frame.pos.astype(np.float64), frame.box, rc)
    achieved = cand * RDF_OPS_PER_CANDIDATE / (rdf_ms * 1e-3) if rdf_ms > 0 else 0.0
    return {"bound": "valu", "candidates_per_frame": cand, "ops_per_candidate": RDF_OPS_PER_CANDIDATE,
            "achieved": achieved, "peak": VALU_PEAK, "unit": "lane-ops/s", "frac": achieved / VALU_PEAK,
            "note": "candidate pairs of the half-stencil cell walk x 7 FP32 operations (3 differences, "
                    "3 multiply-adds, 1 comparison) per second of the RDF stage, against 157.3 TFLOP/s / 2"}


class FrameRun:
    """One context timing K frames of one workload back to back (inputs in HBM)."""

    def __init__(self, wl, frames, precision="fp64", pipeline="serial", charges=None):
        import torch
        from trj_analysis_amd import trjhip as T
        self.T = T
        self.wl, self.frames, self.precision, self.charges = wl, frames, precision, charges
        self.n = n = len(frames[0].itype)
        dev = torch.device("cuda:0")
        self.pos_dtype = T.TRJ_F32 if frames[0].pos.dtype == np.float32 else T.TRJ_F64
        self.dpos = [torch.from_numpy(np.ascontiguousarray(f.pos)).to(dev) for f in frames]
        self.dtyp = [torch.from_numpy(f.itype.astype(np.int32)).to(dev) for f in frames]
        self.dq = self.qh = None
        if charges == "atoms":  # species charges +-0.8 with a small per-atom spread, seeded
            rng = np.random.default_rng(7)
            self.qh = np.where(frames[0].itype == 1, 0.8, -0.8) + 0.05 * rng.standard_normal(n)
            self.dq = torch.from_numpy(self.qh).to(dev)
        elif charges == "species":  # per-type charges
            self.qh = np.where(frames[0].itype == 1, 0.8, -0.8)
        torch.cuda.synchronize()
        self.ctx = ctx = T.Context(0)
        ctx.set_pipeline(T.TRJ_PIPE_OVERLAP if pipeline == "overlap" else T.TRJ_PIPE_SERIAL)
        box = frames[0].box
        rflags = T.TRJ_RDF_CHARGE if charges else 0
        sflags = T.TRJ_SQ_CHARGE if charges else 0
        self.st = ctx.rdf_init(T.rdf_params(3, wl.nsp, wl.deltar, wl.rcrdf, flags=rflags), box)
        qset = T.TRJ_QSET_HALFSPACE if wl.qset == "halfspace" else T.TRJ_QSET_REF
        self.info = ctx.sq_init(T.sq_params(3, wl.nsp, wl.qmax, wl.qmin, wl.bsc, qset, wl.grid, precision,
                                            flags=sflags), box)
        self.nvec = self.info.nvec_grid + self.info.nvec_axis
        self.hpos = ([np.ascontiguousarray(f.pos, dtype=np.float64) for f in frames]
                     if charges == "species" else None)

    def submit(self, k):
        T, ctx, f = self.T, self.ctx, self.frames[k % len(self.frames)]
        if self.charges == "species":
            ctx.frame(self.hpos[k % len(self.frames)], f.itype.astype(np.int32), f.box, charge=self.qh)
            return
        ctx.frame_device(self.dpos[k % len(self.frames)].data_ptr(), self.pos_dtype, T.TRJ_AOS,
                         self.dtyp[k % len(self.frames)].data_ptr(), self.n, f.box,
                         charge_ptr=None if self.dq is None else self.dq.data_ptr())

    def timed(self, steps, warmup, latency_frames=0) -> dict:
        import torch
        ctx = self.ctx
        log(f"warmup {warmup} frames (N={self.n}, lsmax={self.st.lsmax}, q-vectors={self.nvec}, "
            f"{self.precision})")
        for k in range(warmup):
            self.submit(k)
        ctx.sync()
        pairs0 = int(ctx.rdf_counts().sum())
        ctx.timing_reset()
        torch.cuda.synchronize()
        log(f"timed {steps} frames")
        t0 = time.perf_counter()
        for k in range(steps):
            self.submit(warmup + k)
        ctx.sync()
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        pairs = int(ctx.rdf_counts().sum()) - pairs0
        tot, nfr = ctx.timing_totals()
        kernel_ms = {"ingest": tot[3] / nfr, "rdf": tot[0] / nfr, "sq_contract": tot[1] / nfr,
                     "sq_shells_axis": tot[2] / nfr}
        log(f"elapsed {elapsed:.3f} s; kernel ms {kernel_ms}")
        lat = []
        for k in range(latency_frames):
            s0 = time.perf_counter()
            self.submit(k)
            ctx.sync()
            lat.append((time.perf_counter() - s0) * 1e3)
        # With per-atom charges every atom enters the contraction twice: in its
        # species channel and in the charge channel.
        n_entries = 2 * self.n if self.charges == "atoms" else self.n
        flops = 8.0 * n_entries * self.info.nvec_grid
        # the autotuners' choices on the first frame (cell side, LDS tile and
        # waves per cell of the RDF; K-split and workgroup order of the S(Q))
        autotune = {"rdf": ctx.rdf_tuning(), "rdf_cells": ctx.rdf_cell_grid(), "sq": ctx.sq_tuning()}
        return {"steps": steps, "elapsed": elapsed, "fps": steps / elapsed, "pairs": pairs,
                "kernel_ms": kernel_ms, "lat": lat, "flops": flops, "n_entries": n_entries, "autotune": autotune}

    def close(self):
        self.ctx.close()


def shells_of(wl, frame, precision):
    """S(Q) shell sums and RDF counts of one frame on a fresh context."""
    from trj_analysis_amd import trjhip as T
    with T.Context(0) as ctx:
        ctx.rdf_init(T.rdf_params(3, wl.nsp, wl.deltar, wl.rcrdf), frame.box)
        qset = T.TRJ_QSET_HALFSPACE if wl.qset == "halfspace" else T.TRJ_QSET_REF
        ctx.sq_init(T.sq_params(3, wl.nsp, wl.qmax, wl.qmin, wl.bsc, qset, wl.grid, precision), frame.box)
        ctx.frame(frame.pos, frame.itype, frame.box)
        return ctx.sq_shells()[0], ctx.rdf_counts()


def contraction_kernel(precision):
    return {"fp64": "k_sq_contract_f64", "fp32": "k_sq_contract_f32"}.get(
        precision, f"k_sq_contract_lp16w<{precision}>")


def bench_north_star(steps: int, warmup: int) -> dict:
    """The 1M-atom frame (3 species, density 0.85, rc = 5, the 64^3 half-space
    q-grid): frames/s of the plain FP64 per-frame path and of FP16x3 (serial
    and with the RDF overlapped), the 16-bit contraction against its split
    peak, and FP16x3's max / median relative error per |Q| shell against FP64
    on one frame."""
    from trj_analysis_amd import trjhip as T
    from trj_analysis_amd import workloads as W
    wl = W.UNIFORM_1M
    pool = min(steps + warmup, 12)
    log(f"generating {pool} frames of uniform_1m")
    frames = [W.uniform_1m(seed=2 + k) for k in range(pool)]
    runs = {}
    for name, prec, pipe in (("fp64_serial", "fp64", "serial"), ("fp16x3_serial", "fp16x3", "serial"),
                             ("fp16x3_overlap", "fp16x3", "overlap")):
        fr = FrameRun(wl, frames, prec, pipe)
        r = fr.timed(steps, warmup)
        fr.close()
        runs[name] = {"frames_per_s": r["fps"], "ms_per_frame": 1e3 / r["fps"], "kernel_ms": r["kernel_ms"],
                      "autotune": r["autotune"]}
        if pipe == "serial":
            ach = r["flops"] / (r["kernel_ms"]["sq_contract"] * 1e-3)
            runs[name]["contraction"] = {"kernel": contraction_kernel(prec), "achieved": ach / 1e12,
                                         "peak": PEAK[prec] / 1e12, "unit": "TFLOP/s", "frac": ach / PEAK[prec]}
            flops = r["flops"]
    runs["fp16x3_serial"]["rdf_valu"] = rdf_valu_roofline(frames[0], wl.rcrdf,
                                                          runs["fp16x3_serial"]["kernel_ms"]["rdf"])
    base = runs["fp64_serial"]["frames_per_s"]
    for name in runs:
        runs[name]["vs_fp64_serial"] = runs[name]["frames_per_s"] / base
    rate, mhz = T.mfma_rate("fp16x3", 0)
    c16 = runs["fp16x3_serial"]["contraction"]
    c16["measured_mfma"] = {"mfma_rate": rate / 1e12, "clock_mhz": mhz,
                            "frac_of_ceiling": c16["achieved"] * 1e12 / (rate / 3)}
    # Shader cycles per 16-bit MFMA on one SIMD with 2 or 4 FP32 FMAs per MFMA
    # (alone, in the other wavefront of the SIMD, interleaved in the same one):
    # what the contraction's VALU work (B operand, splits, phases) costs beside
    # its MFMAs.
    co = {"mfma_only": T.coissue_rate(1, "mfma")}
    for kv in (2, 4):
        co["valu%d" % kv] = {m: T.coissue_rate(kv, m) for m in ("valu", "cross", "interleaved")}
    c16["coissue_cycles_per_mfma"] = co
    s64, n64 = shells_of(wl, frames[0], "fp64")
    s16, n16 = shells_of(wl, frames[0], "fp16x3")
    m = s64 > 0
    rel = np.abs(s16[m] - s64[m]) / s64[m]
    return {"workload": "uniform_1m", "n_atoms": len(frames[0].itype), "nsp": wl.nsp, "rcrdf": wl.rcrdf,
            "deltar": wl.deltar, "q_grid": list(wl.grid), "q_vectors": int(flops / (8 * len(frames[0].itype))),
            "steps": steps, "warmup": warmup, "paths": runs,
            "speedup_fp16x3": max(runs["fp16x3_serial"]["vs_fp64_serial"], runs["fp16x3_overlap"]["vs_fp64_serial"]),
            "fp16x3_vs_fp64_shell_rel_err": {"max": float(rel.max()), "median": float(np.median(rel)),
                                             "shells": int(m.sum())},
            "rdf_counts_equal": bool(np.array_equal(n64, n16))}


def bench_configs0(cpu: bool) -> dict:
    """The reference's CPU case: 10 frames of a 2-species fluid, N = 4000,
    density 0.8, float32 uniform positions from seed 0, deltar 0.02, rcrdf 4,
    qmax 10 (qmin = qmax, the reference q-set), bsc (1, 0.5).  On the GPU
    (inputs in HBM) and, in full, through the CPU port on the host cores (the
    OpenMP linked-cell RDF and FP64 direct sums over the same q-set); counts and
    shells of the two compared."""
    from oracle import oracle as O
    from trj_analysis_amd import trjhip as T
    from trj_analysis_amd import workloads as W
    wl = W.LJ_BINARY_4K
    frames = list(W.lj_binary_4k(nframes=10))
    fr = FrameRun(wl, frames, "fp64", "serial")
    fr.timed(10, 2)  # warm, then the same 10 frames on a fresh accumulation
    fr.ctx.reset()
    r = fr.timed(10, 0)
    gcounts = fr.ctx.rdf_counts()
    gsqf = fr.ctx.sq_shells()[0]
    fr.close()
    out = {"frames": 10, "n_atoms": 4000, "nsp": 2, "deltar": wl.deltar, "rcrdf": wl.rcrdf, "qmax": wl.qmax,
           "gpu": {"frames_per_s": r["fps"], "ms_per_frame": 1e3 / r["fps"], "kernel_ms": r["kernel_ms"],
                   "precision": "fp64", "inputs": "HBM-resident float32"}}
    if not cpu:
        return out
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    box = frames[0].box
    g = O.sq_setup(3, box, wl.qmax, wl.qmax)
    t_rdf = t_sq = 0.0
    counts = None
    sqf = sqfp = None
    for f in frames:
        t0 = time.perf_counter()
        c = O.rdf_counts(f.pos, f.itype, 3, 2, box, (1, 1, 1), wl.rcrdf, wl.deltar, lsmax=gcounts.shape[1])
        t_rdf += time.perf_counter() - t0
        counts = c if counts is None else counts + c
        pos = O.wrap(f.pos, box, (1, 1, 1))
        t0 = time.perf_counter()
        sqf, sqfp = O.sq_frame_f64(pos, f.itype, 3, 2, box, g, wl.bsc, sqf, sqfp)
        t_sq += time.perf_counter() - t0
    mm = sqf > 0
    out["cpu"] = {"frames_per_s": 10 / (t_rdf + t_sq), "rdf_s": t_rdf, "sq_s": t_sq, "cores": cores,
                  "kind": "port",
                  "what": "OpenMP linked-cell pair histogram (FP64 rule) + FP64 direct sums (double sincos, "
                          "OpenMP over vectors) over the reference q-set; freud is not installed"}
    out["gpu_over_cpu"] = out["gpu"]["frames_per_s"] / out["cpu"]["frames_per_s"]
    out["counts_bit_exact"] = bool(np.array_equal(counts, gcounts))
    out["sq_max_rel_diff"] = float(np.max(np.abs(gsqf[mm] - sqf[mm]) / sqf[mm]))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="lattice_262k",
                    choices=["lattice_262k", "lj_binary_4k", "uniform_1m", "uneven_batch", "serving"])
    ap.add_argument("--batch-frames", type=int, default=512)
    ap.add_argument("--serve-rates", type=float, nargs="+", default=[80.0, 240.0])
    ap.add_argument("--serve-requests", type=int, default=150)
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--precision", default="fp64", choices=list(PEAK))
    ap.add_argument("--pipeline", default="serial", choices=["serial", "overlap"],
                    help="overlap: the RDF runs on its own stream beside the S(Q) contraction")
    ap.add_argument("--latency-frames", type=int, default=10)
    ap.add_argument("--charges", nargs="?", const="atoms", default=None, choices=["atoms", "species"],
                    help="also accumulate g_qq and S_qq: 'atoms' = per-atom charges from HBM "
                         "(trj_frame_device_q); 'species' = per-type charges through the host entry "
                         "trj_frame_q, which forms both from the species sums (inputs host-staged)")
    ap.add_argument("--cpu-baseline-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--extras", default="north_star,configs0,batch,serving",
                    help="comma list of extra objects on the lattice_262k line (empty for none)")
    ap.add_argument("--extra-batch-frames", type=int, default=64)
    ap.add_argument("--extra-serve-requests", type=int, default=40)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "; run it with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.workload == "uneven_batch":
        r = bench_batch(args.batch_frames)
        out = {"metric": METRIC, "value": r["frames_per_s"], "unit": "frames/s", "n_gpus": 1,
               "steps": args.batch_frames, "warmup": 1, "ms_per_step": r["ms_per_frame"],
               "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
               "data": "synthetic", "config": {"workload": "uneven_batch"}, "batch": r}
        print(json.dumps(out), flush=True)
        return
    if args.workload == "serving":
        r = bench_serving(args.serve_rates, args.serve_requests, args.serve_slots)
        value = r["rates"][str(args.serve_rates[-1])]["server"]["throughput_req_s"]
        out = {"metric": METRIC, "value": value, "unit": "requests/s", "n_gpus": 1,
               "steps": args.serve_requests, "warmup": 4, "ms_per_step": 1e3 / value, "higher_is_better": True,
               "scaling": "weak", "vs_baseline": None, "dtype": "fp16x3", "data": "synthetic",
               "config": {"workload": "serving"}, "serving": r}
        print(json.dumps(out), flush=True)
        return

    from trj_analysis_amd import trjhip as T

    pool = min(args.steps + args.warmup, 16)
    log(f"generating {pool} frames of {args.workload}")
    wl, frames = make_frames(args.workload, pool)
    if args.workload != "lj_binary_4k":
        for f in frames:
            f.pos = np.ascontiguousarray(f.pos, dtype=np.float64)
    fr = FrameRun(wl, frames, args.precision, args.pipeline, args.charges)
    r = fr.timed(args.steps, args.warmup, args.latency_frames)
    fr.close()
    n, st, info, nvec = fr.n, fr.st, fr.info, fr.nvec
    kernel_ms, lat, flops = r["kernel_ms"], r["lat"], r["flops"]
    kname = contraction_kernel(args.precision)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get(f"{args.workload}/{args.precision}/{kname}")
    achieved = flops / (kernel_ms["sq_contract"] * 1e-3)
    peak = PEAK[args.precision]
    # The same roofline against the rate the MFMA instruction reaches on this
    # device (trj_mfma_rate: independent accumulations, no memory traffic),
    # scaled like PEAK, and the sustained shader clock over that probe.
    probe = None
    if args.precision in NOMINAL_MFMA:
        rate, mhz = T.mfma_rate(args.precision, 0)
        ceil = rate * peak / NOMINAL_MFMA[args.precision]
        probe = {"mfma_rate": rate / 1e12, "clock_mhz": mhz, "ceiling": ceil / 1e12,
                 "frac_of_ceiling": achieved / ceil}
        if args.precision == "fp64":
            probe["note"] = ("achieved counts 8 flop per complex term; the kernel executes 6 of them on the "
                             "MFMA pipe (three real products per complex product) and skips empty blocks, so "
                             "frac_of_ceiling can pass 1; the MFMA busy fraction from counters is in profiles/")
    # RDF: in-range pairs per second of the RDF stage against the device's
    # measured LDS-atomic rate (one histogram add per in-range pair).
    lds_rate = T.lds_atomic_rate(0)
    pairs_frame = r["pairs"] / args.steps
    rdf_rate = pairs_frame / (kernel_ms["rdf"] * 1e-3) if kernel_ms["rdf"] > 0 else 0.0
    elapsed = r["elapsed"]
    out = {
        "metric": METRIC,
        "value": r["fps"],
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else args.precision,
        "data": "synthetic",
        "config": {"workload": args.workload, "n_atoms": n, "nsp": wl.nsp, "rcrdf": wl.rcrdf,
                   "deltar": wl.deltar, "qmax": wl.qmax, "qset": wl.qset, "q_vectors": nvec,
                   "lsmax": st.lsmax, "precision": args.precision, "pipeline": args.pipeline,
                   "charges": args.charges or "none",
                   "inputs": "host-staged float64" if args.charges == "species" else
                   ("HBM-resident " + ("float32" if fr.pos_dtype == T.TRJ_F32 else "float64"))},
        "pairs_per_s": r["pairs"] / elapsed,
        "q_atom_terms_per_s": nvec * n * args.steps / elapsed,
        "latency_ms_p50": float(np.percentile(lat, 50)) if lat else None,
        "latency_ms_p99": float(np.percentile(lat, 99)) if lat else None,
        "kernel_ms": kernel_ms,
        "autotune": r["autotune"],
        "roofline": {"bound": "mfma", "kernel": kname, "achieved": achieved / 1e12,
                     "peak": peak / 1e12, "unit": "TFLOP/s", "frac": achieved / peak, "traffic": traffic,
                     "algorithmic": f"8*N*Nq = {flops:.4e} flop per launch (N={r['n_entries']} entries, "
                                    f"Nq={info.nvec_grid})",
                     "measured_mfma": probe},
        "rdf_valu": rdf_valu_roofline(frames[0], wl.rcrdf, kernel_ms["rdf"]),
        "rdf_roofline": {"bound": "lds_atomic", "achieved": rdf_rate, "peak": lds_rate, "unit": "pairs/s",
                         "frac": rdf_rate / lds_rate, "pairs_per_frame": pairs_frame,
                         "note": "in-range pairs of the RDF stage (cell build + pairs + merge) per second "
                                 "against the measured LDS-atomic rate (trj_lds_atomic_rate); the pair "
                                 "kernel is bound by its candidate loop (VALU), not by the atomics"},
    }
    default_line = args.workload == "lattice_262k" and args.precision == "fp64" and not args.charges
    extras = [e for e in args.extras.split(",") if e] if default_line else []
    if "north_star" in extras:
        out["north_star"] = bench_north_star(10, 2)
    if "configs0" in extras:
        out["configs0"] = bench_configs0(cpu=not args.no_cpu_baseline)
    if "batch" in extras:
        out["batch"] = bench_batch(args.extra_batch_frames)
    if "serving" in extras:
        out["serving"] = bench_serving([80.0, 240.0], args.extra_serve_requests, args.serve_slots)
    if not args.no_cpu_baseline:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(wl, frames[0], r["pairs"] / max(args.steps, 1),
                                           nvec * n, args.cpu_baseline_seconds)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
