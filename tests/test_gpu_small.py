"""Single sets of at most 256 points run as one kernel (k_small: sort,
dedupe, frame orders, the cell and the extraction in one wavefront) instead
of the general pipeline.  Both must give the oracle's triangles, statuses and
duplicate representatives; the stats show which path ran (a set of at most
256 points is one cell and has no kd level).  Every case prints its name
before it runs."""
import numpy as np
import pytest

import libqdt_amd as Q
from libqdt_amd import workloads as W
import pyoracle as O

pytestmark = pytest.mark.gpu


def run(pts, device=False):
    import torch

    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 2)
    with Q.Triangulation() as t:
        if device:
            d = torch.from_numpy(pts).to("cuda:0")
            st = t.execute_device(d.data_ptr(), len(pts))
        else:
            st = t.execute(pts)
        if st != Q.QDT_OK:
            return st, None, None, None, t.triangles_size()
        used = np.array([t.point_is_used(i) for i in range(len(pts))], dtype=np.uint8)
        return st, t.triangles(), used, t.stats(), t.triangles_size()


def check(name, pts, device=False, exact=None):
    print("case:", name, len(pts), "device" if device else "host", flush=True)
    st, got, used, stats, _ = run(pts, device)
    assert st == Q.QDT_OK, (name, Q.status_string(st))
    want = O.triangulate(pts)
    assert len(got) == len(want.tris), (name, len(got), len(want.tris))
    assert np.array_equal(O.canonical(got), O.canonical(want.tris)), name
    exp = (want.rep == np.arange(len(pts), dtype=np.uint32)).astype(np.uint8) if len(want.tris) else \
        np.zeros(len(pts), np.uint8)
    assert np.array_equal(used, exp), name
    if len(pts) <= 256:
        assert stats["kd_levels"] == 0 and stats["cells"] == 1, (name, stats)  # the one-kernel path ran
    else:
        assert stats["kd_levels"] > 0, (name, stats)
    if exact is not None:
        assert (stats["cells_exact"] == 1) == exact, (name, stats["cells_exact"])
    return stats


def test_sizes_across_the_boundary(gpu):
    for n in (1, 2, 3, 4, 5, 17, 64, 65, 128, 200, 255, 256, 257, 300):
        check("uniform %d" % n, W.uniform(n, 100 + n))


def test_degenerate_small_sets(gpu):
    g = np.array([(i, j) for i in range(16) for j in range(16)], np.float64)
    check("16 x 16 lattice (cocircular: exact cell)", g, exact=True)
    check("lattice, device input", g, device=True, exact=True)
    rng = np.random.default_rng(5)
    dup = np.repeat(rng.random((40, 2)), 5, axis=0)
    rng.shuffle(dup)
    check("40 points, 5 copies each", dup)
    check("one point 200 times", np.tile([[0.25, 0.75]], (200, 1)))
    line = np.stack([np.arange(256.0), 2 * np.arange(256.0) + 1], axis=1)
    check("256 collinear points", line)
    circ = np.stack([np.cos(np.arange(200) * 2 * np.pi / 200), np.sin(np.arange(200) * 2 * np.pi / 200)], axis=1)
    check("200 points on a circle", circ)
    check("-0.0 and +0.0", np.array([[0.0, 0.0], [-0.0, 0.0], [1, 0], [0, 1], [0.0, -0.0]]))
    check("tiny coordinates", W.uniform(150, 3) * 2.0 ** -1000)
    check("huge coordinates", (W.uniform(150, 4) * 2 - 1) * 1.7976931348623157e308)


def test_non_finite_device_input(gpu):
    pts = W.uniform(100, 9)
    pts[37, 1] = np.inf
    print("case: infinite coordinate, device input", flush=True)
    st, _, _, _, T = run(pts, device=True)
    assert st == Q.QDT_ERR_NOT_FINITE
    assert T == 0


def test_verifier_and_mesh_on_small_results(gpu):
    pts = W.uniform(256, 11)
    print("case: verify and mesh of a small result", flush=True)
    with Q.Triangulation() as t:
        assert t.execute(pts) == Q.QDT_OK
        assert t.stats()["kd_levels"] == 0
        v = t.verify()
        assert v["ok"], v
        he, vert, o2l, info = t.mesh()
        assert info["unmatched"] == 0 and info["halfedges"] == 3 * t.triangles_size()
        assert len(vert) == 256
