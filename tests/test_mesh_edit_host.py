"""The library's host-side mesh editor (libqdt_amd/csrc/mesh_edit.cpp, built
on its own as tests/native/libmeshedit.so) against the compiled reference's
qdt_insert_point / qdt_insert_constraint (libqdt/GuibasStolfiDwyer_Wrapper.cpp
:543-876, :1157-1244) on the same operations.

Both start from the same batch triangulation (the oracle's, identical to the
reference's).  With points in general position the (constrained) Delaunay
triangulation is unique, so the canonical triangle lists must be identical
after every step; on lattices (cocircular and collinear points) both must be
valid constrained Delaunay triangulations with the same counts, and every
status code and duplicate representative must be the same."""
import ctypes
import os

import numpy as np
import pytest

import pyoracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "native", "libmeshedit.so")

pytestmark = [pytest.mark.skipif(not os.path.exists(SO), reason="host mesh editor build missing"),
              pytest.mark.skipif(not O.ref_available(), reason="compiled reference not present")]

_L = None


def L():
    global _L
    if _L is None:
        lib = ctypes.CDLL(SO)
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        lib.me_create.restype = vp
        lib.me_destroy.argtypes = [vp]
        lib.me_build.restype = ctypes.c_int
        lib.me_build.argtypes = [vp, vp, sz, vp, vp, sz]
        lib.me_insert.restype = ctypes.c_int
        lib.me_insert.argtypes = [vp, ctypes.c_double, ctypes.c_double, ctypes.POINTER(sz)]
        lib.me_constrain.restype = ctypes.c_int
        lib.me_constrain.argtypes = [vp, sz, sz]
        lib.me_triangles.restype = sz
        lib.me_triangles.argtypes = [vp]
        lib.me_points.restype = sz
        lib.me_points.argtypes = [vp]
        lib.me_copy.argtypes = [vp, vp, vp]
        lib.me_constrained_edges.restype = sz
        lib.me_constrained_edges.argtypes = [vp]
        _L = lib
    return _L


class Editor:
    def __init__(self, pts):
        self.pts = [tuple(p) for p in np.asarray(pts, np.float64)]
        r = O.triangulate(np.asarray(pts, np.float64))
        xy = np.ascontiguousarray(pts, np.float64)
        tris = np.ascontiguousarray(r.tris, np.uint32)
        rep = np.ascontiguousarray(r.rep, np.uint32)
        self.h = L().me_create()
        assert L().me_build(self.h, xy.ctypes.data, len(xy), rep.ctypes.data, tris.ctypes.data, len(tris)) == 1

    def close(self):
        L().me_destroy(self.h)

    def insert(self, x, y):
        idx = ctypes.c_size_t(0)
        st = L().me_insert(self.h, float(x), float(y), ctypes.byref(idx))
        if st == 0:
            self.pts.append((float(x), float(y)))
        return st, idx.value

    def constrain(self, a, b):
        return L().me_constrain(self.h, int(a), int(b))

    def triangles(self):
        t = np.zeros((L().me_triangles(self.h), 3), np.uint32)
        L().me_copy(self.h, t.ctypes.data if len(t) else None, None)
        return t

    def rep(self):
        r = np.zeros(L().me_points(self.h), np.uint32)
        L().me_copy(self.h, None, r.ctypes.data)
        return r


def check_cdt(pts, tris, constraints, rep=None):
    """Exact check of a constrained Delaunay triangulation: triangles strictly
    counter-clockwise, every edge in one or two triangles with opposite
    directions, every accepted constraint (its end points taken through rep)
    present as a chain of edges through the vertices that lie exactly on it
    (points inserted on a constraint split it), every other interior edge
    locally Delaunay, and the Euler count T = 2V - 2 - b."""
    P = np.asarray(pts, np.float64)
    edges = {}
    for t in tris:
        a, b, c = (int(x) for x in t)
        assert O.orient2d(P[a], P[b], P[c]) > 0
        for u, v, w in ((a, b, c), (b, c, a), (c, a, b)):
            assert (u, v) not in edges
            edges[(u, v)] = w
    verts = sorted({int(x) for x in np.asarray(tris).ravel()})
    cons = set()
    for a, b in constraints:
        if rep is not None:
            a, b = int(rep[a]), int(rep[b])
        on = [v for v in verts if v not in (a, b) and O.orient2d(P[a], P[b], P[v]) == 0
              and min(P[a][0], P[b][0]) <= P[v][0] <= max(P[a][0], P[b][0])
              and min(P[a][1], P[b][1]) <= P[v][1] <= max(P[a][1], P[b][1])]
        d = P[b] - P[a]
        chain = [a] + sorted(on, key=lambda v: float(np.dot(P[v] - P[a], d))) + [b]
        for u, v in zip(chain, chain[1:]):
            assert (u, v) in edges or (v, u) in edges, ("constraint missing", a, b, u, v)
            cons.add((min(u, v), max(u, v)))
    boundary = 0
    for (u, v), w in edges.items():
        if (v, u) not in edges:
            boundary += 1
            continue
        if u < v and (u, v) not in cons:
            d = edges[(v, u)]
            assert O.incircle(P[u], P[v], P[w], P[d]) <= 0, ("not locally Delaunay", u, v)
    assert len(tris) == 2 * len(verts) - 2 - boundary


def same_or_valid(ed, rc, general, constraints):
    got, want = ed.triangles(), rc.triangles()
    assert len(got) == len(want)
    if general:
        assert np.array_equal(O.canonical(got), O.canonical(want))
    else:
        check_cdt(ed.pts, got, constraints, ed.rep())


def test_insertions_match_reference_general_position():
    rng = np.random.default_rng(1)
    base = rng.random((3000, 2))
    ed = Editor(base)
    with O.RefContext() as rc:
        assert rc.execute(base) == 0
        extra = np.concatenate([rng.random((600, 2)), rng.random((200, 2)) * 3 - 1])  # inside and outside
        for i, (x, y) in enumerate(extra):
            s1, i1 = ed.insert(x, y)
            s2, i2 = rc.insert(x, y)
            assert (s1, i1) == (s2, i2)
            if i % 100 == 99:
                same_or_valid(ed, rc, True, [])
        same_or_valid(ed, rc, True, [])
    ed.close()


def test_insertions_lattice_duplicates_and_edges():
    """Integer lattice points: many land on edges, on hull edges and on
    existing vertices (duplicates); representatives and statuses equal the
    reference's, and the result is a valid Delaunay triangulation."""
    rng = np.random.default_rng(2)
    base = rng.integers(0, 12, size=(60, 2)).astype(np.float64)
    ed = Editor(base)
    with O.RefContext() as rc:
        assert rc.execute(base) == 0
        for i in range(400):
            x, y = rng.integers(-3, 16, size=2).astype(np.float64)
            s1, i1 = ed.insert(x, y)
            s2, i2 = rc.insert(x, y)
            assert (s1, i1) == (s2, i2)
        same_or_valid(ed, rc, False, [])
        rep = ed.rep()
        used = rc.used()
        assert np.array_equal((rep == np.arange(len(rep))).astype(np.uint8), used)
    ed.close()


def test_constraints_match_reference_general_position():
    rng = np.random.default_rng(3)
    base = rng.random((1500, 2))
    ed = Editor(base)
    accepted = []
    with O.RefContext() as rc:
        assert rc.execute(base) == 0
        for i in range(300):
            a, b = (int(x) for x in rng.integers(0, len(base), size=2))
            if i % 3 == 0:  # short segments: fewer rejections
                d = ((base - base[a]) ** 2).sum(axis=1)
                b = int(np.argsort(d)[int(rng.integers(3, 40))])
            s1 = ed.constrain(a, b)
            s2 = rc.constrain(a, b)
            assert s1 == s2, (i, a, b, s1, s2)
            if s1 == 0:
                accepted.append((a, b))
            if i % 50 == 49:
                same_or_valid(ed, rc, True, accepted)
        assert len(accepted) > 50
        same_or_valid(ed, rc, True, accepted)
        check_cdt(ed.pts, ed.triangles(), accepted, ed.rep())
    ed.close()


def test_constraints_then_insertions():
    """Constraints are kept through later insertions; points inserted exactly
    on a constrained edge split it into two constrained halves."""
    rng = np.random.default_rng(4)
    base = rng.integers(0, 200, size=(400, 2)).astype(np.float64) * 2  # even coordinates
    ed = Editor(base)
    accepted = []
    with O.RefContext() as rc:
        assert rc.execute(base) == 0
        for i in range(200):
            a, b = (int(x) for x in rng.integers(0, len(base), size=2))
            d = ((base - base[a]) ** 2).sum(axis=1)
            b = int(np.argsort(d)[int(rng.integers(2, 12))])
            s1, s2 = ed.constrain(a, b), rc.constrain(a, b)
            assert s1 == s2, (i, a, b, s1, s2)
            if s1 == 0:
                accepted.append((a, b))
        # split constrained edges at their midpoints (exact: even coordinates)
        for a, b in accepted[:60]:
            mx, my = (ed.pts[a][0] + ed.pts[b][0]) / 2, (ed.pts[a][1] + ed.pts[b][1]) / 2
            s1, i1 = ed.insert(mx, my)
            s2, i2 = rc.insert(mx, my)
            assert (s1, i1) == (s2, i2)
        for _ in range(150):
            x, y = rng.integers(0, 400, size=2).astype(np.float64)
            assert ed.insert(x, y) == rc.insert(x, y)
        got, want = ed.triangles(), rc.triangles()
        assert len(got) == len(want)
        check_cdt(ed.pts, got, accepted, ed.rep())
    ed.close()


def test_constraint_rejections():
    """Through a point, crossing a constraint, same point, bad index: the same
    codes as the reference, and nothing changes on a rejection."""
    pts = np.array([(i, j) for i in range(6) for j in range(6)], np.float64)
    ed = Editor(pts)
    with O.RefContext() as rc:
        assert rc.execute(pts) == 0
        seq = [(0, 35), (0, 14), (6, 1), (0, 7), (1, 6), (2, 2), (0, 36), (5, 30), (3, 8), (3, 20), (12, 25),
               (13, 24), (0, 5)]
        for a, b in seq:
            before = O.canonical(ed.triangles())
            s1, s2 = ed.constrain(a, b), rc.constrain(a, b)
            assert s1 == s2, (a, b, s1, s2)
            if s1 != 0:
                assert np.array_equal(before, O.canonical(ed.triangles()))
    ed.close()
