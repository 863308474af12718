"""The device half-edge mesh (qdt_gpu_build_mesh) against the oracle's
restatement of ensure_mesh_ (GuibasStolfiDwyer_Wrapper.cpp:99-202).

On the GPU path's own triangle list the device arrays must equal the
oracle's bit for bit (origin, twin, next, triangle per half-edge; input index
and first half-edge per local vertex; the input-to-local map), on random,
clustered, duplicated, lattice, collinear and tiny inputs.  At 2M points the
mesh is checked through properties (mutual reversed twins, boundary count
equal to the verifier's, one local vertex per distinct point).  An index out
of range is rejected.  Every case prints its name before it runs."""
import numpy as np
import pytest

import libqdt_amd as Q
from libqdt_amd import workloads as W
import pyoracle as O

pytestmark = pytest.mark.gpu


def dev(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


CASES = [
    ("uniform 20k", lambda: W.uniform(20000, 21)),
    ("gaussian 20k", lambda: W.gaussian(20000, 22)),
    ("clustered 20k", lambda: W.clustered(20000, 23)),
    ("lattice+dups+circle", lambda: W.lattice_dups_circle(side=60, ndup=800, ncircle=1200, seed=4)),
    ("grid 25x25", lambda: np.array([(i, j) for i in range(25) for j in range(25)], dtype=np.float64)),
    ("collinear 40", lambda: np.array([(i, 3.0 * i) for i in range(40)], dtype=np.float64)),
    ("three points", lambda: np.array([(0.0, 0.0), (1.0, 0.0), (0.0, 1.0)])),
]


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_mesh_matches_oracle(gpu, name, make):
    pts = make()
    print("case:", name, len(pts), flush=True)
    with Q.Triangulation() as t:
        assert t.execute(pts) == Q.QDT_OK
        tris = t.triangles()
        he, vert, o2l, info = t.mesh()
    want_he, want_vert, want_o2l = O.build_mesh(tris, len(pts))
    assert info["halfedges"] == 3 * len(tris)
    assert info["unmatched"] == 0
    assert info["vertices"] == len(want_vert)
    assert info["boundary"] == int((want_he[:, 1] < 0).sum())
    assert np.array_equal(he, want_he), name
    assert np.array_equal(vert, want_vert), name
    assert np.array_equal(o2l, want_o2l), name


def test_mesh_bad_index(gpu):
    print("case: index out of range", flush=True)
    import torch

    d_tris = dev(np.array([0, 1, 2, 0, 2, 7], dtype=np.uint32).view(np.int32))
    he = torch.empty((6, 4), dtype=torch.int32, device="cuda")
    vert = torch.empty((6, 2), dtype=torch.int32, device="cuda")
    o2l = torch.empty(5, dtype=torch.int32, device="cuda")
    with pytest.raises(RuntimeError, match="index"):
        Q.build_mesh(d_tris.data_ptr(), 2, 5, he.data_ptr(), vert.data_ptr(), o2l.data_ptr())


def test_mesh_large_properties(gpu):
    import torch

    n = 2_000_000
    print("case: uniform", n, flush=True)
    pts = W.uniform(n, 24)
    d = dev(pts)
    with Q.Triangulation() as t:
        assert t.execute_device(d.data_ptr(), n) == Q.QDT_OK
        r = t.verify()
        assert r["ok"]
        d_tris, T = t.device_triangles()
        he = torch.empty((3 * T, 4), dtype=torch.int32, device="cuda")
        vert = torch.empty((min(n, 3 * T), 2), dtype=torch.int32, device="cuda")
        o2l = torch.empty(n, dtype=torch.int32, device="cuda")
        info = Q.build_mesh(d_tris, T, n, he.data_ptr(), vert.data_ptr(), o2l.data_ptr())
        torch.cuda.synchronize()
        print("mesh", info, flush=True)
        assert info["unmatched"] == 0
        assert info["boundary"] == r["boundary_edges"]
        assert info["vertices"] == r["vertices"] == t.stats()["unique_points"]
        h = torch.arange(3 * T, device="cuda", dtype=torch.int64)
        tw = he[:, 1].long()
        has = tw >= 0
        assert int((~has).sum()) == info["boundary"]
        assert bool((he[tw[has], 1].long() == h[has]).all())
        nxt = he[:, 2].long()
        assert bool((he[tw[has], 0] == he[nxt[has], 0]).all())
        assert bool((nxt == 3 * (h // 3) + (h % 3 + 1) % 3).all())
        v = vert[: info["vertices"]]
        assert bool((o2l[v[:, 0].long()] == torch.arange(info["vertices"], device="cuda", dtype=torch.int32)).all())
        # first half-edges ascend with the local id (order of first appearance)
        assert bool((v[1:, 1] > v[:-1, 1]).all())
