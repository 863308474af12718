"""The oracle's half-edge mesh (restating ensure_mesh_,
GuibasStolfiDwyer_Wrapper.cpp:99-202), pinned on the CPU.

A two-triangle example is written out by hand from the definition; on
oracle triangulations of random, clustered and duplicated inputs the mesh is
checked against properties fixed by the definition and by geometry: half-edge
3t+i starts at corner i of triangle t and is followed by 3t+(i+1)%3; twins
are mutual and run the other way; the half-edges without a twin are exactly
the boundary edges (as many as the boundary vertices the exact checker
counts); local vertices are the distinct indices used, in order of first
appearance (numpy's unique with first indices), each with its first
half-edge."""
import numpy as np
import pytest

import pyoracle as O
from libqdt_amd import workloads as W


def test_two_triangles_by_hand():
    # unit square split along 0-2: (0, 1, 2) and (0, 2, 3); index 4 unused
    he, vert, o2l = O.build_mesh(np.array([[0, 1, 2], [0, 2, 3]]), 5)
    # half-edges 0:0->1 1:1->2 2:2->0 3:0->2 4:2->3 5:3->0; 2 and 3 are twins
    assert he.tolist() == [[0, -1, 1, 0], [1, -1, 2, 0], [2, 3, 0, 0],
                           [0, 2, 4, 1], [2, -1, 5, 1], [3, -1, 3, 1]]
    assert vert.tolist() == [[0, 0], [1, 1], [2, 2], [3, 5]]
    assert o2l.tolist() == [0, 1, 2, 3, -1]


def test_first_appearance_order_by_hand():
    # vertex 7 appears first, then 3, then 5, then 9
    he, vert, o2l = O.build_mesh(np.array([[7, 3, 5], [5, 3, 9]]), 10)
    assert vert[:, 0].tolist() == [7, 3, 5, 9]
    assert vert[:, 1].tolist() == [0, 1, 2, 5]
    assert he[:, 0].tolist() == [0, 1, 2, 2, 1, 3]
    # shared edge 3-5: half-edge 1 (3->5) and 3 (5->3)
    assert he[1, 1] == 3 and he[3, 1] == 1
    assert o2l[7] == 0 and o2l[3] == 1 and o2l[5] == 2 and o2l[9] == 3 and o2l[0] == -1


def test_empty_and_bad_index():
    he, vert, o2l = O.build_mesh(np.zeros((0, 3), np.uint32), 4)
    assert he.shape == (0, 4) and len(vert) == 0 and o2l.tolist() == [-1] * 4
    with pytest.raises(ValueError):
        O.build_mesh(np.array([[0, 1, 4]]), 4)


def check_mesh_properties(pts, tris, he, vert, o2l):
    T = len(tris)
    flat = tris.reshape(-1)
    h = np.arange(3 * T)
    assert (he[:, 2] == 3 * (h // 3) + (h % 3 + 1) % 3).all()
    assert (he[:, 3] == h // 3).all()
    assert (he[:, 0] == o2l[flat]).all()
    tw = he[:, 1]
    has = tw >= 0
    # mutual, reversed
    assert (tw[tw[has]] == h[has]).all()
    assert (he[tw[has], 0] == he[he[has, 2], 0]).all()
    assert (he[he[tw[has], 2], 0] == he[has, 0]).all()
    # no twin exactly on the boundary: as many as the boundary vertices
    rc, msg, info = O.check(pts, tris)
    assert rc == 0, msg
    assert int((~has).sum()) == info["b"]
    # local vertices: distinct used indices in order of first appearance
    u, first = np.unique(flat, return_index=True)
    order = np.argsort(first)
    assert vert[:, 0].tolist() == u[order].tolist()
    assert vert[:, 1].tolist() == first[order].tolist()
    assert (o2l[vert[:, 0]] == np.arange(len(vert))).all()
    assert int((o2l >= 0).sum()) == len(vert)


CASES = [
    ("uniform 3000", lambda: W.uniform(3000, 11)),
    ("gaussian 3000", lambda: W.gaussian(3000, 12)),
    ("clustered 3000", lambda: W.clustered(3000, 13)),
    ("lattice+dups+circle", lambda: W.lattice_dups_circle(side=30, ndup=300, ncircle=400, seed=3)),
]


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_oracle_mesh_properties(name, make):
    pts = make()
    r = O.triangulate(pts)
    he, vert, o2l = O.build_mesh(r.tris, len(pts))
    assert len(vert) == r.unique
    check_mesh_properties(pts, r.tris, he, vert, o2l)
