"""The exact device verifier (qdt_gpu_verify, qdt_gpu_verify_triangles).

Correct results of the GPU path pass on every input kind (random, clustered,
lattice with duplicates and cocircular points, collinear, tiny); lists with a
mistake planted on the host (a reversed triangle, a flipped diagonal, a
dropped or repeated triangle, an index out of range, a duplicate's
non-lowest index) fail with the matching count.  The largest input a context
accepts (100M points) passes too.  Every case prints its name before it
runs."""
import numpy as np
import pytest

import libqdt_amd as Q
from libqdt_amd import workloads as W
import pyoracle as O

pytestmark = pytest.mark.gpu


def dev(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def check_list(pts, tris, rep=None):
    d_pts = dev(np.asarray(pts, dtype=np.float64).reshape(-1, 2))
    d_tris = dev(np.asarray(tris, dtype=np.uint32).reshape(-1).view(np.int32))
    d_rep = dev(np.asarray(rep, dtype=np.uint32).view(np.int32)) if rep is not None else None
    return Q.verify_triangles(d_pts.data_ptr(), len(d_pts), d_tris.data_ptr(), len(tris),
                              d_rep.data_ptr() if d_rep is not None else None)


CASES = [
    ("uniform 50k", lambda: W.uniform(50000, 3)),
    ("gaussian 50k", lambda: W.gaussian(50000, 4)),
    ("clustered 50k", lambda: W.clustered(50000, 5)),
    ("lattice+dups+circle small", lambda: W.lattice_dups_circle(side=100, ndup=2000, ncircle=3000, seed=2)),
    ("grid 30x30", lambda: np.array([(i, j) for i in range(30) for j in range(30)], dtype=np.float64)),
    ("collinear 50", lambda: np.array([(i, 2.0 * i) for i in range(50)], dtype=np.float64)),
    ("three points", lambda: np.array([(0.0, 0.0), (1.0, 0.0), (0.0, 1.0)])),
]


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_gpu_results_pass(gpu, name, make):
    pts = make()
    print("case:", name, len(pts), flush=True)
    want = O.triangulate(pts)
    for device in (False, True):
        with Q.Triangulation() as t:
            if device:
                d = dev(pts)
                assert t.execute_device(d.data_ptr(), len(pts)) == Q.QDT_OK
            else:
                assert t.execute(pts) == Q.QDT_OK
            r = t.verify()
            assert r["ok"], (name, device, r)
            assert r["triangles"] == len(want.tris)
            if len(want.tris):
                assert r["vertices"] == len(np.unique(want.rep))
                assert r["triangles"] + 2 + r["boundary_edges"] == 2 * r["vertices"]


def test_planted_mistakes_fail(gpu):
    pts = W.uniform(20000, 9)
    print("case: planted mistakes, uniform 20k", flush=True)
    res = O.triangulate(pts)
    tris = res.tris.astype(np.uint32).copy()
    base = check_list(pts, tris, res.rep)
    assert base["ok"], base

    bad = tris.copy()
    bad[10, [1, 2]] = bad[10, [2, 1]]  # reversed
    r = check_list(pts, bad, res.rep)
    assert not r["ok"] and r["bad_orientation"] == 1

    # flip an interior diagonal: (a, b, c) and (b, a, d) -> (a, d, c) and (d, b, c)
    edge_tri = {}
    for t, (a, b, c) in enumerate(tris.tolist()):
        for u, v in ((a, b), (b, c), (c, a)):
            edge_tri[(u, v)] = t
    flipped = None
    for t, (a, b, c) in enumerate(tris.tolist()):
        t2 = edge_tri.get((b, a))
        if t2 is None:
            continue
        d = [x for x in tris[t2].tolist() if x not in (a, b)][0]
        P = pts
        o1 = O.orient2d(P[a], P[d], P[c])
        o2 = O.orient2d(P[d], P[b], P[c])
        if o1 > 0 and o2 > 0:
            flipped = (t, t2, [a, d, c], [d, b, c])
            break
    assert flipped is not None
    bad = tris.copy()
    bad[flipped[0]] = flipped[2]
    bad[flipped[1]] = flipped[3]
    r = check_list(pts, bad, res.rep)
    assert not r["ok"] and r["non_delaunay"] >= 1 and r["bad_orientation"] == 0 and r["bad_edges"] == 0, r

    r = check_list(pts, np.delete(tris, 500, axis=0), res.rep)  # a hole
    assert not r["ok"], r

    r = check_list(pts, np.concatenate([tris, tris[:1]]), res.rep)  # a repeated triangle
    assert not r["ok"] and r["bad_edges"] >= 1, r

    bad = tris.copy()
    bad[3, 0] = len(pts) + 5  # out of range
    r = check_list(pts, bad, res.rep)
    assert not r["ok"] and r["bad_index"] == 1, r

    r = check_list(pts, tris[:0], res.rep)  # no triangles for points that are not collinear
    assert not r["ok"] and r["off_line"] > 0, r


def test_duplicate_index_fails(gpu):
    pts = np.array([(0.0, 0.0), (1.0, 0.0), (0.0, 1.0), (1.0, 1.0), (1.0, 0.0)])  # 4 duplicates 1
    print("case: non-representative index", flush=True)
    res = O.triangulate(pts)
    assert check_list(pts, res.tris, res.rep)["ok"]
    bad = res.tris.astype(np.uint32).copy()
    bad[bad == 1] = 4
    r = check_list(pts, bad, res.rep)
    assert not r["ok"] and r["not_representative"] >= 1 and r["unused_points"] == 1, r


def test_large_inputs_pass(gpu):
    # 10M uniform (the bench input) and the 5.2M degenerate mix: lattice,
    # duplicates and a jittered circle, where most predicates go exact
    for name, make in (("uniform 10M", lambda: W.uniform(10_000_000, 1)),
                       ("lattice 2000^2 + 200k duplicates + 1M circle", lambda: W.lattice_dups_circle())):
        pts = make()
        print("case:", name, len(pts), flush=True)
        with Q.Triangulation() as t:
            d = dev(pts)
            assert t.execute_device(d.data_ptr(), len(pts)) == Q.QDT_OK
            r = t.verify()
            st = t.stats()
        assert r["ok"], (name, r)
        assert r["vertices"] == st["unique_points"] and r["triangles"] == st["triangles"]


def test_maximum_size_passes(gpu):
    # the largest input a context accepts (wrapper.hxx:15, kMaxPoints =
    # 100,000,000): the whole path at that size, checked exactly on the device
    n = 100_000_000
    print("case: uniform", n, flush=True)
    import torch

    pts = W.uniform(n, 7)
    d = dev(pts)
    del pts
    with Q.Triangulation() as t:
        assert t.execute_device(d.data_ptr(), n) == Q.QDT_OK
        r = t.verify()
        st = t.stats()
    del d
    torch.cuda.empty_cache()
    assert r["ok"], r
    assert st["unique_points"] == n and r["vertices"] == n
    # Euler with b boundary vertices: T = 2n - 2 - b
    assert r["triangles"] == st["triangles"] == 2 * n - 2 - r["boundary_edges"]
