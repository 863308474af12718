"""Batched form (qdt_gpu_batch_*): every set of a batch gets exactly the
triangles it gets alone.

Each set is checked against the oracle (canonical triangle list, distinct
count, duplicate representatives) and, for the large mixed batch, against the
single-instance GPU path, so a result is shown to be the same alone and in
the mix.  Every case prints its name before it runs.
"""
import numpy as np
import pytest

import libqdt_amd as Q
from libqdt_amd import workloads as W
import pyoracle as O

pytestmark = pytest.mark.gpu


def concat(sets):
    offs = np.zeros(len(sets) + 1, dtype=np.uint64)
    for i, p in enumerate(sets):
        offs[i + 1] = offs[i] + len(p)
    pts = np.concatenate([np.asarray(p, dtype=np.float64).reshape(-1, 2) for p in sets]) if sets else \
        np.zeros((0, 2))
    return np.ascontiguousarray(pts), offs


def run_batch(sets, device=False):
    pts, offs = concat(sets)
    b = Q.Batch()
    if device:
        import torch

        d = torch.from_numpy(pts).to("cuda:0")
        st = b.execute_device(d.data_ptr(), offs)
    else:
        st = b.execute(pts, offs)
    assert st == Q.QDT_OK, Q.status_string(st)
    return b, pts, offs


def check_sets(name, sets, device=False):
    print("case:", name, len(sets), "sets", sum(len(p) for p in sets), "points", flush=True)
    b, pts, offs = run_batch(sets, device)
    allt = b.all_triangles()
    rep = b.rep(len(pts))
    assert b.sets() == len(sets)
    for i, p in enumerate(sets):
        p = np.asarray(p, dtype=np.float64).reshape(-1, 2)
        finite = bool(np.isfinite(p).all())
        if not finite:
            assert b.status(i) == Q.QDT_ERR_NOT_FINITE, (name, i)
            assert b.triangles_size(i) == 0
            continue
        assert b.status(i) == Q.QDT_OK, (name, i)
        want = O.triangulate(p)
        got = b.triangles(i, allt)
        assert len(got) == b.triangles_size(i)
        assert len(got) == len(want.tris), (name, i, len(p), len(got), len(want.tris))
        assert np.array_equal(O.canonical(got), O.canonical(want.tris)), (name, i)
        assert b.unique_points(i) == want.unique, (name, i)
        assert np.array_equal(rep[int(offs[i]):int(offs[i + 1])], want.rep), (name, i)
        if len(got):
            t0 = b.get_triangle(i, 0)
            assert t0 == tuple(int(x) for x in got[0])
    return b


def test_small_mixed_sets(gpu):
    rng = np.random.default_rng(11)
    sets = []
    for n in [0, 1, 2, 3, 4, 5, 7, 10, 33, 100, 257, 1000, 4000]:
        sets.append(rng.random((n, 2)))
        sets.append(rng.integers(-3, 4, size=(n, 2)).astype(np.float64))  # duplicates and ties
    sets.append([(i, 2 * i) for i in range(40)])                      # collinear
    sets.append([(5.0, 5.0)] * 50)                                      # one distinct point
    sets.append([(i, j) for i in range(30) for j in range(30)])        # lattice
    bad = rng.random((100, 2))
    bad[17, 1] = np.nan
    sets.insert(5, bad)                                                 # rejected alone
    inf = rng.random((10, 2))
    inf[3, 0] = np.inf
    sets.append(inf)
    check_sets("small mixed", sets)
    check_sets("small mixed, device input", sets, device=True)


def test_single_set_batch_equals_plain(gpu):
    p = W.uniform(50000, 7)
    b = check_sets("one set of 50k", [p])
    with Q.Triangulation() as t:
        assert t.execute(p) == Q.QDT_OK
        assert np.array_equal(O.canonical(t.triangles()), O.canonical(b.triangles(0)))


def test_batch_api_errors(gpu):
    b = Q.Batch()
    pts = np.zeros((4, 2))
    assert b.execute(pts, [1, 4]) == Q.QDT_ERR_INVALID_INDEX
    assert b.execute(pts, [0, 3, 2]) == Q.QDT_ERR_INVALID_INDEX
    assert b.execute(pts, [0]) == Q.QDT_OK and b.sets() == 0
    assert b.status(0) == Q.QDT_ERR_INVALID_INDEX
    assert b.get_triangle(0, 0) is None


def test_uneven_batch_alone_and_in_the_mix(gpu):
    """256 sets, sizes log-uniform in [64, 65536], the uniform / Gaussian /
    32-cluster mix (the batched workload, scaled down): every set equals the
    oracle and the single-instance GPU path."""
    sets = W.batch_sets(count=256, seed=3)
    b = check_sets("uneven 256", sets)
    allt = b.all_triangles()
    with Q.Triangulation() as t:
        for i in range(0, len(sets), 16):
            assert t.execute(sets[i]) == Q.QDT_OK
            assert np.array_equal(O.canonical(t.triangles()), O.canonical(b.triangles(i, allt))), i
    # the same batch again (triangle order inside a set may differ between
    # runs: records of seam pieces come from an atomic counter)
    b2, _, _ = run_batch(sets)
    allt2 = b2.all_triangles()
    for i in range(len(sets)):
        assert np.array_equal(O.canonical(b2.triangles(i, allt2)), O.canonical(b.triangles(i, allt))), i


def test_copies_without_y_ties(gpu):
    """Sets of random points where some points are repeated (exact copies,
    the same point in two sets) but no two distinct points share a y value:
    the frame-1 order then comes from the y sort grouped by set, and the
    duplicate rule (lowest input index represents) must still hold."""
    rng = np.random.default_rng(11)
    base = rng.random((3000, 2))
    sets = []
    for k in range(12):
        n = int(rng.integers(4, 400))
        p = base[rng.choice(len(base), size=n, replace=True)]  # copies inside a set
        sets.append(p)
    sets.append(sets[0].copy())  # the same points in another set
    check_sets("copies, no y ties", sets)


@pytest.mark.skipif(not O.ref_available(), reason="compiled reference not present")
def test_c4_full_batch_matches_reference(gpu):
    """The batched workload at its own size: 4,096 sets, sizes log-uniform in
    [64, 65536] (seed 3), uniform / Gaussian / 32-cluster mix, in one
    qdt_gpu_batch_execute_device call; every set's triangle count and
    canonical FNV-1a-64 equal the compiled reference's qdt_execute on that
    set alone (one context per host thread)."""
    import torch

    sets = W.batch_sets(count=4096, seed=3)
    pts, offs = concat(sets)
    print("case: C4 4096 sets,", len(pts), "points", flush=True)
    d = torch.from_numpy(pts).to("cuda:0")
    b = Q.Batch()
    assert b.execute_device(d.data_ptr(), offs) == Q.QDT_OK
    assert b.sets() == 4096
    allt = b.all_triangles()
    print("case: C4 reference hashes", flush=True)
    cnt, hs = O.ref_batch_hashes(pts, offs, threads=16)
    bad = []
    for i in range(len(sets)):
        got = b.triangles(i, allt)
        if len(got) != int(cnt[i]) or O.canon_hash(got) != int(hs[i]):
            bad.append(i)
    assert not bad, "sets differing from the reference: %s" % bad[:20]
    assert int(allt.shape[0]) == int(cnt.sum())
    b.close()
