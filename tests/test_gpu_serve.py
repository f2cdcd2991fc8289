"""The request server: results are the same alone and in the mix; admission
respects the memory budget."""
import numpy as np
import pytest

from trj_analysis_amd import serve as S
from trj_analysis_amd import trjhip as T

pytestmark = pytest.mark.gpu


def test_alone_equals_mixed():
    jobs = S.make_jobs(8, rate=1e6, seed=3, nmax=6e4, precision="fp64")
    print("case: serve alone vs mixed", [len(j.frame.itype) for j in jobs],
          [j.frame.periodic for j in jobs], flush=True)
    _, mixed, _ = S.serve_stream(jobs, slots=4)
    for k, j in enumerate(jobs):
        _, alone, _ = S.serve_stream([j], slots=1)
        a, m = alone[0], mixed[k]
        np.testing.assert_array_equal(a[0], m[0])  # counts bit-exact
        for x, y in ((a[1], m[1]), (a[2], m[2])):
            mask = np.abs(x) > 0
            assert np.max(np.abs(x[mask] - y[mask]) / np.abs(x[mask])) < 1e-12
        np.testing.assert_array_equal(a[3], m[3])


def test_sequential_equals_server():
    jobs = S.make_jobs(4, rate=1e6, seed=5, nmax=4e4, precision="fp16x3")
    print("case: serve vs sequential", flush=True)
    _, srv, _ = S.serve_stream(jobs, slots=3)
    _, seq, _ = S.serve_sequential(jobs)
    for a, b in zip(srv, seq):
        np.testing.assert_array_equal(a[0], b[0])
        mask = np.abs(b[1]) > 0
        assert np.max(np.abs(a[1][mask] - b[1][mask]) / np.abs(b[1][mask])) < 1e-12


def test_budget_admission():
    jobs = S.make_jobs(5, rate=1e6, seed=9, nmax=3e4, precision="fp64")
    reqs = [T.request(j.frame.pos, j.frame.itype, j.frame.box, rdf=j.rdf, sq=j.sq) for j in jobs]
    fp = [T.request_footprint(r) for r, _ in reqs]
    print("case: budget", fp, flush=True)
    srv = T.Server(slots=4, budget_bytes=max(fp) + 1)  # room for one request at a time
    tickets = [srv.submit(r, keep) for r, keep in reqs]
    seen = 0
    while True:
        q, running, reserved = srv.stats()
        assert running <= 1 and reserved <= max(fp) + 1
        seen = max(seen, running)
        if q == 0 and running == 0:
            break
    for tk, j in zip(tickets, jobs):
        while not srv.poll(tk):
            pass
        srv.result(tk, j.lsmax, j.nit, j.nqmax, j.frame.nsp)
    assert srv.stats() == (0, 0, 0)
    srv.close()
