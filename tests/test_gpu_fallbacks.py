"""The engine's alternative and fallback paths, each forced from the host and
checked against the oracle: merges without seam pieces, kd levels without LDS
cells, a spare-record region too small for the seam pieces (the run is
repeated without them), sorts split into their top radix passes and a fix-up
whose fix-up meets a long run out of order (the run is repeated with full
sorts), and a device-memory limit that turns the workspace
allocation into QDT_ERR_OUT_OF_MEMORY with an emptied context
(libqdt/qdt.cpp:63-77, 107-112: bad_alloc -> QDT_ERR_OUT_OF_MEMORY, context
reset).  Every case prints its name before it runs."""
import numpy as np
import pytest

import libqdt_amd as Q
from libqdt_amd import workloads as W
import pyoracle as O
import _fixture_inputs

pytestmark = pytest.mark.gpu


def run(pts, seams=True, cells=True, spare=0, mem=0):
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 2)
    with Q.Triangulation() as t:
        t.set_seams(seams)
        t.set_cells(cells)
        t.set_test_limits(mem, spare)
        st = t.execute(pts)
        assert st == Q.QDT_OK, Q.status_string(st)
        return t.triangles(), t.stats()


def same(name, pts, **kw):
    print("case:", name, len(pts), kw, flush=True)
    got, st = run(pts, **kw)
    want = O.triangulate(pts)
    assert len(got) == len(want.tris), name
    assert np.array_equal(O.canonical(got), O.canonical(want.tris)), name
    return st


def inputs():
    d = dict(_fixture_inputs.inputs())
    d["lattice+dups+circle small"] = W.lattice_dups_circle(side=120, ndup=3000, ncircle=5000, seed=2)
    d["gaussian 200k"] = W.gaussian(200000, 3)
    d["clustered 200k"] = W.clustered(200000, 4)
    return d


@pytest.mark.parametrize("seams,cells", [(False, True), (True, False), (False, False)])
def test_switches_give_identical_results(gpu, seams, cells):
    for name, p in inputs().items():
        same("%s seams=%d cells=%d" % (name, seams, cells), p, seams=seams, cells=cells)


def test_seam_retry_when_spare_records_run_out(gpu):
    for name, p in (("uniform 300k", W.uniform(300000, 5)), ("gaussian 300k", W.gaussian(300000, 6)),
                    ("lattice+dups+circle", W.lattice_dups_circle(side=400, ndup=20000, ncircle=40000, seed=7))):
        st = same(name + " default", p)
        assert st["seam_retries"] == 0 and st["seam"]["pieces"] > 0, "seam pieces not used"
        st = same(name + " spare=64", p, spare=64)
        assert st["seam_retries"] == 1, st["seam_retries"]


def near_ties(n, seed, axis):
    """n points whose coordinate on `axis` takes 1,000 values, each moved by a
    random multiple (below 2^16) of its unit in the last place: order keys
    equal in all but their two lowest bytes, in runs of ~1,100 keys in input
    order, far longer than the split sorts' fix-up takes."""
    rng = np.random.default_rng(seed)
    g = 0.5 + rng.integers(0, 1000, n) / 1000.0
    v = g + rng.integers(0, 65536, n) * np.spacing(g)
    p = np.empty((n, 2))
    p[:, axis] = v
    p[:, 1 - axis] = rng.random(n)
    return p


def test_split_sorts_and_the_full_sort_repeat(gpu):
    # a single set of at least 2^20 points sorts on the top five bytes in
    # which its keys differ, then fixes the runs equal in those bytes; a run
    # too long for that, out of order, sets this error bit and the run is
    # repeated with full sorts
    split_err = 1 << 22
    st = same("uniform 1.1M, split sorts", W.uniform(1100000, 9))
    assert not st["err_bits"] & split_err, st["err_bits"]
    for ax, name in ((1, "y"), (0, "x")):
        st = same("near ties in %s, 1.1M" % name, near_ties(1100000, 11 + ax, ax))
        assert st["err_bits"] & split_err, st["err_bits"]


def test_out_of_memory_empties_the_context(gpu):
    p = W.uniform(100000, 8)
    with Q.Triangulation() as t:
        print("case: OOM after a good run", flush=True)
        assert t.execute(p) == Q.QDT_OK and t.triangles_size() > 0
        t.release_workspace()
        t.set_test_limits(1 << 20, 0)
        assert t.execute(p) == Q.QDT_ERR_OUT_OF_MEMORY
        assert t.triangles_size() == 0 and t.points_size() == 0
        assert t.get_triangle(0) is None and t.point_is_used(0) == 0
        print("case: the emptied context is reusable", flush=True)
        t.set_test_limits(0, 0)
        assert t.execute(p) == Q.QDT_OK
        want = O.triangulate(p)
        assert np.array_equal(O.canonical(t.triangles()), O.canonical(want.tris))
        print("case: OOM on device input", flush=True)
        import torch

        d = torch.from_numpy(p).to("cuda:0")
        t.release_workspace()
        t.set_test_limits(1 << 20, 0)
        assert t.execute_device(d.data_ptr(), len(p)) == Q.QDT_ERR_OUT_OF_MEMORY
        assert t.triangles_size() == 0 and t.points_size() == 0
