"""Serving form (qdt_gpu_server_*): requests submitted while others run are
batched together, and every result equals the oracle's for that request
alone.  Every case prints its name before it runs."""
import threading
import time

import numpy as np
import pytest

import libqdt_amd as Q
from libqdt_amd import workloads as W
import pyoracle as O

pytestmark = pytest.mark.gpu


def test_requests_alone_and_in_the_mix(gpu):
    print("case: 60 mixed requests from two threads", flush=True)
    rng = np.random.default_rng(3)
    sizes = [int(x) for x in W.log_uniform_sizes(60, 10, 200000, 9)]
    sets = [W.mixed_set(n, 1000 + i)[0] for i, n in enumerate(sizes)]
    sets[7] = rng.integers(-2, 3, size=(500, 2)).astype(np.float64)  # duplicates and ties
    sets[11] = np.zeros((0, 2))
    bad = rng.random((50, 2))
    bad[4, 0] = np.nan
    sets[13] = bad
    ids = [None] * len(sets)
    with Q.Server(hbm_budget_bytes=8 << 30, workers=2) as srv:
        def feed(lo):
            for i in range(lo, len(sets), 2):
                ids[i] = srv.submit(sets[i])
                time.sleep(0.001 * (i % 3))
        th = [threading.Thread(target=feed, args=(k,)) for k in (0, 1)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        for i, p in enumerate(sets):
            st, lat = srv.wait(ids[i])
            assert lat >= 0
            if i == 13:
                assert st == Q.QDT_ERR_NOT_FINITE
                continue
            assert st == Q.QDT_OK, (i, Q.status_string(st))
            want = O.triangulate(p)
            got = srv.triangles(ids[i])
            assert np.array_equal(O.canonical(got), O.canonical(want.tris)), i
            assert srv.unique_points(ids[i]) == want.unique, i
            srv.release(ids[i])
        batches, reqs = srv.counters()
        assert reqs == len(sets)
        assert batches < len(sets)   # at least some requests shared a batch


def test_budget_rejects_what_cannot_fit(gpu):
    print("case: budget", flush=True)
    # a request is charged ~320 B per point plus 4 MB per batch: 1,000 points
    # fit in 16 MB, 100,000 do not
    with Q.Server(hbm_budget_bytes=16 << 20, workers=1) as srv:
        small = srv.submit(W.uniform(1000, 1))
        big = srv.submit(W.uniform(100000, 2))
        assert srv.wait(small)[0] == Q.QDT_OK
        assert srv.wait(big)[0] == Q.QDT_ERR_OUT_OF_MEMORY
        assert srv.poll(12345) == -1


@pytest.mark.skipif(not O.ref_available(), reason="compiled reference not present")
def test_served_workload_matches_reference(gpu):
    """The served workload's requests at their own sizes: 32 requests of
    log-uniform size in [1,000, 2,000,000] (seed 5, the uniform / Gaussian /
    32-cluster mix), submitted with Poisson arrivals at 500 requests/s into a
    server with a 200 GB budget and 4 workers; each result equals the
    compiled reference's qdt_execute of that request alone."""
    nreq = 32
    sizes = W.log_uniform_sizes(nreq, 1000, 2_000_000, 5)
    reqs = [W.mixed_set(int(sz), 5_000_011 + i)[0] for i, sz in enumerate(sizes)]
    offs = np.zeros(nreq + 1, np.uint64)
    for i, r in enumerate(reqs):
        offs[i + 1] = offs[i] + len(r)
    print("case: served 32 requests,", int(offs[-1]), "points", flush=True)
    cnt, hs = O.ref_batch_hashes(np.concatenate(reqs), offs, threads=16)
    gaps = np.random.default_rng(17).exponential(1.0 / 500, size=nreq)
    with Q.Server(hbm_budget_bytes=200 << 30, workers=4) as srv:
        ids = []
        t0 = time.perf_counter()
        at = 0.0
        for i, r in enumerate(reqs):
            at += gaps[i]
            delay = t0 + at - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            ids.append(srv.submit(r))
        for i, rid in enumerate(ids):
            st, lat = srv.wait(rid)
            assert st == Q.QDT_OK, (i, Q.status_string(st))
            got = srv.triangles(rid)
            assert len(got) == int(cnt[i]), i
            assert O.canon_hash(got) == int(hs[i]), i
            srv.release(rid)
        batches, served = srv.counters()
        assert served == nreq
        held, peak = srv.memory()
        print("served: %d batches, device bytes held %d, peak %d" % (batches, held, peak), flush=True)
        assert 0 < peak <= 200 << 30


def test_budget_bounds_the_workers_memory(gpu):
    """A budget below what four workers would keep: the workers' charges
    (retained workspaces included) never exceed it, idle workers give memory
    back, and every request still completes with the oracle's result."""
    budget = 1 << 30
    sizes = [int(x) for x in W.log_uniform_sizes(24, 50000, 1_200_000, 11)]
    reqs = [W.uniform(n, 7000 + i) for i, n in enumerate(sizes)]
    print("case: budget 1 GiB, 24 requests,", sum(sizes), "points", flush=True)
    with Q.Server(hbm_budget_bytes=budget, workers=4) as srv:
        ids = [srv.submit(r) for r in reqs]
        for i, rid in enumerate(ids):
            st, _ = srv.wait(rid)
            assert st == Q.QDT_OK, (i, Q.status_string(st))
            if i % 6 == 0:
                want = O.triangulate(reqs[i])
                assert np.array_equal(O.canonical(srv.triangles(rid)), O.canonical(want.tris)), i
            srv.release(rid)
        held, peak = srv.memory()
        print("held %d peak %d" % (held, peak), flush=True)
        assert peak <= budget and held <= budget
