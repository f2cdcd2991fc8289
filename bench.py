#!/usr/bin/env python3
"""Benchmark of the structural hot path (RDF + S(Q)) on one MI355X.

    python bench.py [--steps K] [--warmup W] [--workload lattice_262k]

A step is one frame through the whole path: ingest (wrap, species sort), the
cell-list pair histogram g_ab(r) and the S(Q) contraction with its shell sums.
Frames are generated from a seed and copied to HBM before the timed region;
the timed region runs K frames back to back on the context stream and ends
with a device synchronisation.  One JSON line is printed.

Fields besides the driver's contract:
  pairs_per_s         in-range pairs binned per second (RDF)
  q_atom_terms_per_s  (q-vector, atom) phase terms per second (S(Q))
  latency_ms_p50/p99  per-frame latency, frames submitted one at a time
  kernel_ms           mean per-frame time of each stage (HIP events on the
                      context stream)
  roofline            the dominant kernel (the S(Q) contraction) against the
                      FP64 peak: algorithmic flops 8*N*Nq per launch / mean
                      launch time
  cpu_baseline        the CPU oracle on the host cores, on a bounded sample of
                      the same workload, extrapolated to frames/s
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
    # S(Q): long-double direct sums over a sample of the q-set.
    pos = O.wrap(frame.pos, box, per)
    rng = np.random.default_rng(0)
    kv = rng.integers(0, 8, size=(16, 3)).astype(np.int32)
    t0 = time.perf_counter()
    O.sq_rho(pos, frame.itype, wl.nsp, box, kv)
    t16 = time.perf_counter() - t0
    m = int(max(16, min(20000, 16 * max(0.0, budget_s - t_rdf) / max(t16, 1e-6))))
    kv = rng.integers(0, 32, size=(m, 3)).astype(np.int32)
    t0 = time.perf_counter()
    O.sq_rho(pos, frame.itype, wl.nsp, box, kv)
    t_m = time.perf_counter() - t0
    n = len(frame.itype)
    terms_rate = m * n / t_m
    t_sq = nq_terms / terms_rate
    pairs_rate = float(counts.sum()) / t_rdf
    fps = 1.0 / (t_rdf + t_sq)
    return {
        "value": fps, "unit": "frames/s", "cores": cores, "kind": "port",
        "sample": (f"one {n}-atom frame: RDF in full (linked cells, OpenMP, {t_rdf:.2f} s); S(Q) "
                   f"long-double direct sums over {m} q-vectors ({t_m:.2f} s), extrapolated to the "
                   f"{nq_terms // n} vectors of a frame"),
        "pairs_per_s": pairs_rate, "q_atom_terms_per_s": terms_rate,
    }


def bench_batch(args):
    """uneven_batch: frames of N log-uniform in [1e4, 1e6], 2-4 species, float32,
    rcrdf=4, deltar=0.01; per-frame pair histograms.  Timed: the batched
    graph-replayed path (autotuned on the first frames) against the
    one-frame-per-launch path on the same frames; results compared bit for bit."""
    import torch
    from trj_analysis_amd import trjhip as T
    from trj_analysis_amd import workloads as W

    nf = args.batch_frames
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
            ctx.rdf_init(T.rdf_params(3, f.nsp, wl.deltar, wl.rcrdf, f.periodic), f.box)
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
    out = {
        "metric": METRIC, "value": nf / t_batch, "unit": "frames/s", "n_gpus": 1, "steps": nf,
        "warmup": 1, "ms_per_step": t_batch * 1e3 / nf, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "uneven_batch", "frames": nf, "atoms": natoms, "rcrdf": wl.rcrdf,
                   "deltar": wl.deltar, "inputs": "HBM-resident float32",
                   "autotuned": {"frames_per_launch": chosen.frames_per_launch,
                                 "workgroups": chosen.workgroups,
                                 "waves_per_cell": chosen.waves_per_cell}},
        "pairs_per_s": pairs / t_batch,
        "plain_frames_per_s": nf / t_plain,
        "speedup_vs_plain": t_plain / t_batch,
        "no_graph_frames_per_s": nf / t_nograph,
        "batch_gpu_ms": gpu_ms,
        "bit_exact_vs_plain": exact,
    }
    b.close()
    nog.close()
    print(json.dumps(out), flush=True)


def bench_serving(args):
    """serving: Poisson streams of RDF + S(Q) requests (bulk or z-slab frames, N
    log-uniform in [1e4, 2e6], 1-4 species, float32, 64^3 half-space q-grid,
    FP16x3 contraction) at two offered rates; the continuous-batching server
    against the plain per-request path on the same arrivals."""
    from trj_analysis_amd import serve as S
    out_rates = {}
    value = None
    for rate in args.serve_rates:
        jobs = S.make_jobs(args.serve_requests, rate=rate, seed=11, precision="fp16x3")
        log(f"rate {rate}/s: {len(jobs)} requests, mean N {np.mean([len(j.frame.itype) for j in jobs]):.0f}")
        S.serve_stream(jobs[:4], slots=args.serve_slots)  # warm
        lat_s, res_s, span_s = S.serve_stream(jobs, slots=args.serve_slots)
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
        value = len(jobs) / span_s
    out = {"metric": METRIC, "value": value, "unit": "requests/s", "n_gpus": 1,
           "steps": args.serve_requests, "warmup": 4, "ms_per_step": 1e3 / value, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "fp16x3", "data": "synthetic",
           "config": {"workload": "serving", "slots": args.serve_slots, "rates": args.serve_rates,
                      "rcrdf": S.RCRDF, "deltar": S.DELTAR, "q_grid": list(S.GRID)},
           "rates": out_rates}
    print(json.dumps(out), flush=True)


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
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "; run it with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.workload == "uneven_batch":
        bench_batch(args)
        return
    if args.workload == "serving":
        bench_serving(args)
        return

    import torch
    from trj_analysis_amd import trjhip as T

    pool = min(args.steps + args.warmup, 16)
    log(f"generating {pool} frames of {args.workload}")
    wl, frames = make_frames(args.workload, pool)
    n = len(frames[0].itype)
    dev = torch.device("cuda:0")
    dpos = [torch.from_numpy(np.ascontiguousarray(f.pos, dtype=np.float64)).to(dev) for f in frames]
    dtyp = torch.from_numpy(frames[0].itype.astype(np.int32)).to(dev)
    dq = None
    qh = None
    if args.charges == "atoms":  # species charges +-0.8 with a small per-atom spread, seeded
        rng = np.random.default_rng(7)
        qh = np.where(frames[0].itype == 1, 0.8, -0.8) + 0.05 * rng.standard_normal(n)
        dq = torch.from_numpy(qh).to(dev)
    elif args.charges == "species":  # per-type charges
        qh = np.where(frames[0].itype == 1, 0.8, -0.8)
    torch.cuda.synchronize()

    ctx = T.Context(0)
    ctx.set_pipeline(T.TRJ_PIPE_OVERLAP if args.pipeline == "overlap" else T.TRJ_PIPE_SERIAL)
    box = frames[0].box
    rflags = T.TRJ_RDF_CHARGE if args.charges else 0
    sflags = T.TRJ_SQ_CHARGE if args.charges else 0
    st = ctx.rdf_init(T.rdf_params(3, wl.nsp, wl.deltar, wl.rcrdf, flags=rflags), box)
    qset = T.TRJ_QSET_HALFSPACE if wl.qset == "halfspace" else T.TRJ_QSET_REF
    info = ctx.sq_init(T.sq_params(3, wl.nsp, wl.qmax, wl.qmin, wl.bsc, qset, wl.grid, args.precision,
                                   flags=sflags), box)
    nvec = info.nvec_grid + info.nvec_axis

    hpos = [np.ascontiguousarray(f.pos, dtype=np.float64) for f in frames] if args.charges == "species" else None
    hit = frames[0].itype.astype(np.int32)

    def submit(k):
        if args.charges == "species":
            ctx.frame(hpos[k % pool], hit, box, charge=qh)
            return
        ctx.frame_device(dpos[k % pool].data_ptr(), T.TRJ_F64, T.TRJ_AOS, dtyp.data_ptr(), n, box,
                         charge_ptr=None if dq is None else dq.data_ptr())

    log(f"warmup {args.warmup} frames (N={n}, lsmax={st.lsmax}, q-vectors={nvec})")
    for k in range(args.warmup):
        submit(k)
    ctx.sync()
    pairs0 = int(ctx.rdf_counts().sum())
    ctx.timing_reset()
    torch.cuda.synchronize()
    log(f"timed {args.steps} frames")
    t0 = time.perf_counter()
    for k in range(args.steps):
        submit(args.warmup + k)
    ctx.sync()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    pairs = int(ctx.rdf_counts().sum()) - pairs0
    tot, nfr = ctx.timing_totals()
    kernel_ms = {"ingest": tot[3] / nfr, "rdf": tot[0] / nfr, "sq_contract": tot[1] / nfr,
                 "sq_shells_axis": tot[2] / nfr}
    log(f"elapsed {elapsed:.3f} s; kernel ms {kernel_ms}")

    lat = []
    for k in range(args.latency_frames):
        s0 = time.perf_counter()
        submit(k)
        ctx.sync()
        lat.append((time.perf_counter() - s0) * 1e3)

    fps = args.steps / elapsed
    # With charges every atom enters the contraction twice: in its species
    # channel and in the charge channel.
    n_entries = 2 * n if args.charges == "atoms" else n
    flops = 8.0 * n_entries * info.nvec_grid
    kname = {"fp64": "k_sq_contract_f64", "fp32": "k_sq_contract_f32"}.get(args.precision,
                                                                          f"k_sq_contract_lp16w<{args.precision}>")
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
    # RDF: in-range pairs per second of the RDF stage against the device's
    # measured LDS-atomic rate (one histogram add per in-range pair).
    lds_rate = T.lds_atomic_rate(0)
    pairs_frame = pairs / args.steps
    rdf_rate = pairs_frame / (kernel_ms["rdf"] * 1e-3) if kernel_ms["rdf"] > 0 else 0.0
    out = {
        "metric": METRIC,
        "value": fps,
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
                   "inputs": "host-staged float64" if args.charges == "species" else "HBM-resident float64"},
        "pairs_per_s": pairs / elapsed,
        "q_atom_terms_per_s": nvec * n * args.steps / elapsed,
        "latency_ms_p50": float(np.percentile(lat, 50)) if lat else None,
        "latency_ms_p99": float(np.percentile(lat, 99)) if lat else None,
        "kernel_ms": kernel_ms,
        "roofline": {"bound": "mfma", "kernel": kname, "achieved": achieved / 1e12,
                     "peak": peak / 1e12, "unit": "TFLOP/s", "frac": achieved / peak, "traffic": traffic,
                     "algorithmic": f"8*N*Nq = {flops:.4e} flop per launch (N={n_entries} entries, Nq={info.nvec_grid})",
                     "measured_mfma": probe},
        "rdf_roofline": {"bound": "lds_atomic", "achieved": rdf_rate, "peak": lds_rate, "unit": "pairs/s",
                         "frac": rdf_rate / lds_rate, "pairs_per_frame": pairs_frame,
                         "note": "in-range pairs of the RDF stage (cell build + pairs + merge) per second "
                                 "against the measured LDS-atomic rate (trj_lds_atomic_rate); the pair "
                                 "kernel is bound by its candidate loop (VALU), not by the atomics"},
    }
    ctx.close()
    if not args.no_cpu_baseline:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(wl, frames[0], pairs / max(args.steps, 1),
                                           nvec * n, args.cpu_baseline_seconds)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
