"""Poisson request stream over the trj_hip request server.

Each request is one frame (RDF + S(Q)): a 3D bulk box or a z-confined slab
(periodic in x and y only), N log-uniform in [1e4, 2e6], 1-4 species,
float32 uniform positions from a per-request seed (workloads.serving_frame).
Requests arrive at a fixed offered rate (exponential inter-arrival times).

  serve_stream(...)     the continuous-batching server: requests are admitted
                        into free slots as they arrive, their work overlaps
  serve_sequential(...) the plain per-request path: one context, requests run
                        one after the other in arrival order

Latency = completion time - arrival time (queueing included).
"""
from __future__ import annotations

import time
from dataclasses import dataclass
from typing import List

import numpy as np

from . import trjhip as T
from . import workloads as W

RCRDF, DELTAR = 4.0, 0.01
GRID = (64, 64, 64)


@dataclass
class Job:
    frame: W.Frame
    arrival: float
    rdf: T.RdfParams
    sq: T.SqParams
    lsmax: int
    nit: int
    nqmax: int


def make_job(f: W.Frame, arrival: float = 0.0, precision: str = "fp16x3") -> Job:
    """One RDF + S(Q) request on frame f: rc = 4, deltar = 0.01, the 64^3
    half-space q-grid, unit scattering lengths."""
    # one frame per request: no first-frame autotuning of the pair kernels
    rdf = T.rdf_params(3, f.nsp, DELTAR, RCRDF, f.periodic, flags=T.TRJ_RDF_NO_AUTOTUNE)
    sq = T.sq_params(3, f.nsp, 1e9, bsc=[1.0] * f.nsp, qset=T.TRJ_QSET_HALFSPACE, grid=GRID,
                     precision=precision)
    ls, _ = T.rdf_plan(rdf, f.box)
    info, _ = T.sq_plan(sq, f.box)
    return Job(f, arrival, rdf, sq, ls, f.nsp * (f.nsp + 1) // 2, info.nqmax)


def make_jobs(nreq: int, rate: float, seed: int = 7, nmax: float = 2e6, precision: str = "fp16x3") -> List[Job]:
    rng = np.random.default_rng(seed)
    t = 0.0
    jobs = []
    for k in range(nreq):
        t += rng.exponential(1.0 / rate)
        jobs.append(make_job(W.serving_frame(seed * 1000 + k, 1e4, nmax), t, precision))
    return jobs


def serve_stream(jobs: List[Job], slots: int = 4, budget_bytes: int = 0):
    """Continuous batching: returns (latencies_s, results, makespan_s)."""
    srv = T.Server(slots, budget_bytes)
    reqs = [T.request(j.frame.pos, j.frame.itype, j.frame.box, rdf=j.rdf, sq=j.sq) for j in jobs]
    lat = [0.0] * len(jobs)
    res = [None] * len(jobs)
    t0 = time.perf_counter()
    nxt, live = 0, {}
    while nxt < len(jobs) or live:
        now = time.perf_counter() - t0
        while nxt < len(jobs) and jobs[nxt].arrival <= now:
            r, keep = reqs[nxt]
            live[srv.submit(r, keep)] = nxt
            nxt += 1
        for tk in list(live):
            if srv.poll(tk):
                k = live.pop(tk)
                j = jobs[k]
                res[k] = srv.result(tk, j.lsmax, j.nit, j.nqmax, j.frame.nsp)
                lat[k] = (time.perf_counter() - t0) - j.arrival
        if not live and nxt < len(jobs):
            wait = jobs[nxt].arrival - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(min(wait, 0.002))
        else:
            time.sleep(0.0002)
    span = time.perf_counter() - t0
    srv.close()
    return lat, res, span


def serve_sequential(jobs: List[Job]):
    """The plain per-request path: one context, one request at a time."""
    ctx = T.Context(0)
    lat, res = [], []
    t0 = time.perf_counter()
    for j in jobs:
        now = time.perf_counter() - t0
        if now < j.arrival:
            time.sleep(j.arrival - now)
        ctx.rdf_init(j.rdf, j.frame.box)
        ctx.sq_init(j.sq, j.frame.box)
        ctx.frame(j.frame.pos, j.frame.itype, j.frame.box)
        counts = ctx.rdf_counts()
        sqf, sqfp, nq = ctx.sq_shells()
        res.append((counts, sqf, sqfp, nq))
        lat.append((time.perf_counter() - t0) - j.arrival)
    span = time.perf_counter() - t0
    ctx.close()
    return lat, res, span
