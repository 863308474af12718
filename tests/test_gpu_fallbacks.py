"""The engine's alternative and fallback paths, each forced from the host and
checked against the oracle: merges without seam pieces, kd levels without LDS
cells, a spare-record region too small for the seam pieces (the run is
repeated without them), and a device-memory limit that turns the workspace
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
