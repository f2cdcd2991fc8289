"""The request server: results are the same alone and in the mix, agree with
the oracle at the serving workload's full size, and the device memory the
slots hold never exceeds the HBM budget."""
import numpy as np
import pytest

from oracle import oracle as O
from trj_analysis_amd import serve as S
from trj_analysis_amd import trjhip as T
from trj_analysis_amd import workloads as W

pytestmark = pytest.mark.gpu


def _frame(seed, n, nsp, slab):
    """A serving frame with a chosen size: float32 uniform positions at density
    0.85, random species fractions, a z-confined slab (periodic in x, y) or a
    bulk box."""
    rng = np.random.Generator(np.random.PCG64(seed))
    fr = rng.dirichlet(np.ones(nsp))
    counts = np.maximum(1, np.floor(fr * n).astype(int))
    counts[-1] = n - counts[:-1].sum()
    itype = np.concatenate([np.full(c, s + 1, dtype=np.int32) for s, c in enumerate(counts)])
    rng.shuffle(itype)
    L = (n / 0.85) ** (1.0 / 3.0)
    pos = (rng.random((n, 3)) * L).astype(np.float32)
    return W.Frame(pos, itype, np.array([L, L, L]), nsp, periodic=(True, True, not slab), seed=seed)


def _same(a, m, tol=0.0):
    np.testing.assert_array_equal(a[0], m[0])  # counts bit-exact
    for x, y in ((a[1], m[1]), (a[2], m[2])):
        mask = np.abs(x) > 0
        assert np.max(np.abs(x[mask] - y[mask]) / np.abs(x[mask])) <= tol
    np.testing.assert_array_equal(a[3], m[3])


def test_alone_equals_mixed():
    jobs = S.make_jobs(8, rate=1e6, seed=3, nmax=6e4, precision="fp64")
    print("case: serve alone vs mixed", [len(j.frame.itype) for j in jobs],
          [j.frame.periodic for j in jobs], flush=True)
    _, mixed, _ = S.serve_stream(jobs, slots=4)
    for k, j in enumerate(jobs):
        _, alone, _ = S.serve_stream([j], slots=1)
        _same(alone[0], mixed[k], 1e-12)


def test_sequential_equals_server():
    jobs = S.make_jobs(4, rate=1e6, seed=5, nmax=4e4, precision="fp16x3")
    print("case: serve vs sequential", flush=True)
    _, srv, _ = S.serve_stream(jobs, slots=3)
    _, seq, _ = S.serve_sequential(jobs)
    for a, b in zip(srv, seq):
        np.testing.assert_array_equal(a[0], b[0])
        mask = np.abs(b[1]) > 0
        assert np.max(np.abs(a[1][mask] - b[1][mask]) / np.abs(b[1][mask])) < 1e-12


def test_full_size_mix_against_oracle():
    """The serving workload at its stated size: a 1.6M-atom z-confined slab and
    bulk frames (1e4 - 2e5 atoms, 1-4 species), RDF + FP16x3 S(Q) on the 64^3
    half-space grid, submitted together to a 4-slot server.
      (i)  each request's result in the mix equals the same request alone
           (counts bit-exact; S(Q) shells to 1e-12: the FP64 shell atomics may
           reorder);
      (ii) the slab's RDF counts are bit-exact against the oracle's cell lists;
      (iii) 100 rho vectors of the slab, read from a context running the same
           request, agree with the oracle's long-double sums (FP16x3 bound:
           2e-5 * sqrt(N_s) per species channel), and that context's shells
           equal the server's."""
    frames = [_frame(101, 1_600_000, 3, True), _frame(102, 200_000, 2, False), _frame(103, 10_000, 1, False),
              _frame(104, 50_000, 4, True)]
    jobs = [S.make_job(f, 0.0, "fp16x3") for f in frames]
    print("case: serve full size", [len(f.itype) for f in frames], [f.periodic for f in frames], flush=True)
    _, mixed, _ = S.serve_stream(jobs, slots=4)
    for k, j in enumerate(jobs):
        _, alone, _ = S.serve_stream([j], slots=1)
        _same(alone[0], mixed[k], 1e-12)
    print("alone == mixed", flush=True)
    big, jb = frames[0], jobs[0]
    want = O.rdf_counts(big.pos, big.itype, 3, 3, big.box, big.periodic, S.RCRDF, S.DELTAR, lsmax=jb.lsmax,
                        method="cells")
    np.testing.assert_array_equal(mixed[0][0], want)
    print("slab counts bit-exact, pairs", int(want.sum()), flush=True)
    rng = np.random.default_rng(5)
    kv = np.stack([rng.integers(-32, 32, 110), rng.integers(-32, 32, 110), rng.integers(0, 32, 110)], 1)
    kv = kv[np.any(kv != 0, axis=1)][:100].astype(np.int32)
    with T.Context(0) as ctx:
        ctx.rdf_init(jb.rdf, big.box)
        ctx.sq_init(jb.sq, big.box)
        ctx.frame(big.pos, big.itype, big.box)
        rho = ctx.sq_rho(kv)
        sqf, sqfp, nq = ctx.sq_shells()
    _same((mixed[0][0], sqf, sqfp, nq), mixed[0], 1e-12)
    orho = O.sq_rho(O.wrap(big.pos, big.box, big.periodic), big.itype, 3, big.box, kv)
    ns = np.bincount(big.itype, minlength=4)[1:]
    err = np.abs(rho - orho) / np.sqrt(ns)[None, :]
    print("slab rho: max |drho|/sqrt(N_s) %.3e, max |rho| %.3e" % (err.max(), np.abs(orho).max()), flush=True)
    assert err.max() <= 2e-5


def test_footprint_bounds_fresh_context():
    """trj_request_footprint is an upper bound on what a new context allocates
    for the request (bulk and slab, 1-4 species, FP64 and FP16x3 S(Q), the
    reference octant q-set and the half-space grid)."""
    cases = [(_frame(201, 20_000, 1, False), "fp16x3"), (_frame(202, 300_000, 4, True), "fp16x3"),
             (_frame(203, 60_000, 2, False), "fp64"), (_frame(204, 1_000_000, 3, False), "fp16x3")]
    for f, prec in cases:
        j = S.make_job(f, 0.0, prec)
        r, _ = T.request(f.pos, f.itype, f.box, rdf=j.rdf, sq=j.sq)
        fp = T.request_footprint(r)
        with T.Context(0) as ctx:
            ctx.rdf_init(j.rdf, f.box)
            ctx.sq_init(j.sq, f.box)
            ctx.frame(f.pos, f.itype, f.box)
            ctx.sync()
            used = ctx.device_bytes()
        print("case: footprint", len(f.itype), f.nsp, f.periodic, prec, "used", used, "bound", fp,
              "ratio %.3f" % (fp / used), flush=True)
        assert used <= fp
    # the reference octant q-set (partially filled tiles on the sphere)
    f = _frame(205, 40_000, 2, False)
    rdf = T.rdf_params(3, 2, 0.01, 4.0)
    sq = T.sq_params(3, 2, 6.0, 6.0, [1.0, 0.5])
    r, _ = T.request(f.pos, f.itype, f.box, rdf=rdf, sq=sq)
    fp = T.request_footprint(r)
    with T.Context(0) as ctx:
        ctx.rdf_init(rdf, f.box)
        ctx.sq_init(sq, f.box)
        ctx.frame(f.pos, f.itype, f.box)
        ctx.sync()
        used = ctx.device_bytes()
    print("case: footprint ref q-set used", used, "bound", fp, flush=True)
    assert used <= fp


def test_budget_admission():
    """Budget for one request at a time: never more than one running, the
    slots' retained buffers plus the running footprint never above the budget,
    and no request outgrows what was reserved for it."""
    jobs = S.make_jobs(5, rate=1e6, seed=9, nmax=3e4, precision="fp64")
    reqs = [T.request(j.frame.pos, j.frame.itype, j.frame.box, rdf=j.rdf, sq=j.sq) for j in jobs]
    fp = [T.request_footprint(r) for r, _ in reqs]
    budget = max(fp) + 1
    print("case: budget", fp, flush=True)
    srv = T.Server(slots=4, budget_bytes=budget)  # room for one request at a time
    tickets = [srv.submit(r, keep) for r, keep in reqs]
    while True:
        q, running, reserved = srv.stats()
        assert running <= 1 and reserved <= budget
        if q == 0 and running == 0:
            break
    res = []
    for tk, j in zip(tickets, jobs):
        while not srv.poll(tk):
            pass
        res.append(srv.result(tk, j.lsmax, j.nit, j.nqmax, j.frame.nsp))
    reserved, peak, misses = srv.memory()
    print("reserved", reserved, "peak", peak, "budget", budget, "misses", misses, flush=True)
    assert peak <= budget and misses == 0
    assert srv.stats()[:2] == (0, 0)
    srv.close()
    _, seq, _ = S.serve_sequential(jobs)
    for a, b in zip(res, seq):
        _same(a, b, 1e-12)


def test_over_budget_request_refused():
    f = _frame(301, 50_000, 2, False)
    j = S.make_job(f, 0.0, "fp64")
    r, keep = T.request(f.pos, f.itype, f.box, rdf=j.rdf, sq=j.sq)
    fp = T.request_footprint(r)
    print("case: over budget", fp, flush=True)
    srv = T.Server(slots=2, budget_bytes=fp // 2)
    with pytest.raises(T.TrjError) as e:
        srv.submit(r, keep)
    assert e.value.code == T.TRJ_ELIMIT
    srv.close()
