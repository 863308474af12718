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
    with Q.Server(hbm_budget_bytes=1 << 20, workers=1) as srv:
        small = srv.submit(W.uniform(1000, 1))
        big = srv.submit(W.uniform(100000, 2))
        assert srv.wait(small)[0] == Q.QDT_OK
        assert srv.wait(big)[0] == Q.QDT_ERR_OUT_OF_MEMORY
        assert srv.poll(12345) == -1
