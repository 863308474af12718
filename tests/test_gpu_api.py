"""libqdt's API contract on the GPU build (restated from test/qdt_verify.cpp
scenarios api, guards, examples and contexts, :745-1031 and :2052-2102)."""
import threading

import numpy as np
import pytest

import libqdt_amd as Q
from libqdt_amd import workloads as W
import pyoracle as O

pytestmark = pytest.mark.gpu

OK4 = np.array([[0, 0], [4, 0], [3, 1], [1, 1]], np.float64)


def test_guards_empty_the_context(gpu):
    for par in (False, True):
        with Q.Triangulation() as t:
            print("guards parallel=%s" % par, flush=True)
            assert t.execute(OK4, par) == Q.QDT_OK and t.triangles_size() == 2
            assert Q.lib().qdt_execute(t.handle, None, 4) == Q.QDT_ERR_NULL_ARGUMENT
            assert t.triangles_size() == 0 and t.points_size() == 0
            for bad in ([np.nan, 1], [np.inf, 1], [1, -np.inf]):
                assert t.execute(OK4, par) == Q.QDT_OK
                p = OK4.copy()
                p[2] = bad
                assert t.execute(p, par) == Q.QDT_ERR_NOT_FINITE
                assert t.triangles_size() == 0 and t.points_size() == 0
            assert t.execute(OK4, par) == Q.QDT_OK
            assert Q.lib().qdt_execute(t.handle, OK4.ctypes.data, 100000001) == Q.QDT_ERR_TOO_MANY_POINTS
            assert t.triangles_size() == 0 and t.points_size() == 0
            assert t.execute(OK4, par) == Q.QDT_OK and t.triangles_size() == 2


def test_queries_and_copy_semantics(gpu):
    with Q.Triangulation() as t:
        d = OK4.copy()
        assert t.execute(d) == Q.QDT_OK
        d[:] = -999.0
        assert t.triangles_size() == 2 and t.points_size() == 4
        assert t.get_point(2) == (3.0, 1.0) and t.get_point(4) is None
        assert all(t.point_is_used(i) == 1 for i in range(4)) and t.point_is_used(4) == 0
        tri = t.get_triangle(0)
        assert tri is not None and t.get_triangle(2) is None
        assert O.orient2d(OK4[tri[0]], OK4[tri[1]], OK4[tri[2]]) > 0
        dup = np.array([[0, 0], [1, 0], [0, 0], [0, 1], [1, 0]], np.float64)
        assert t.execute(dup) == Q.QDT_OK
        assert [t.point_is_used(i) for i in range(5)] == [1, 1, 0, 1, 0]
        assert t.execute(dup[:2]) == Q.QDT_OK
        assert t.points_size() == 2 and t.point_is_used(0) == 0 and t.triangles_size() == 0
        line = np.array([[0, 0], [1, 1], [2, 2]], np.float64)
        assert t.execute(line) == Q.QDT_OK and t.triangles_size() == 0
        assert t.insert_constraint(0, 2) == Q.QDT_ERR_NO_TRIANGLES


def test_insert_rebuilds_delaunay(gpu):
    # README six-point example: square, centre, and a point on an edge
    pts = np.array([[0, 0], [1, 0], [0, 1], [1, 1], [0.5, 0.5], [0.25, 0.75]], np.float64)
    with Q.Triangulation() as t:
        assert t.execute(pts[:4]) == Q.QDT_OK
        assert t.insert_point(0.5, 0.5) == (Q.QDT_OK, 4)
        assert t.insert_point(0.25, 0.75) == (Q.QDT_OK, 5)
        got = t.triangles()
        assert len(got) == 6
        want = np.array([[1, 3, 4], [0, 1, 4], [2, 0, 5], [0, 4, 5], [4, 3, 5], [3, 2, 5]])
        assert O.canonical(got).tolist() == O.canonical(want).tolist()
    with Q.Triangulation() as t:
        assert t.execute(np.zeros((0, 2))) == Q.QDT_OK
        for i in range(5):
            assert t.insert_point(float(i), float(i * i))[0] == Q.QDT_OK
        assert t.triangles_size() == 3


def test_constraint_rules(gpu):
    with Q.Triangulation() as t:
        assert t.execute(OK4) == Q.QDT_OK
        assert t.insert_constraint(0, 99) == Q.QDT_ERR_INVALID_INDEX
        assert t.insert_constraint(2, 2) == Q.QDT_ERR_SAME_POINT
        # 1-3 is the Delaunay diagonal of the quad (README.md:326-331): already an edge
        assert t.insert_constraint(1, 3) == Q.QDT_OK
        assert t.insert_constraint(3, 1) == Q.QDT_OK
        assert t.insert_constraint(0, 2) == Q.QDT_ERR_CROSSES_CONSTRAINT


def test_independent_contexts_in_threads(gpu):
    sets = [W.uniform(50000 + 1000 * i, 100 + i) for i in range(8)]
    want = [O.canon_hash(O.triangulate(p).tris) for p in sets]
    got = [None] * 8
    errs = []

    def work(i):
        try:
            with Q.Triangulation() as t:
                for _ in range(2):
                    assert t.execute(sets[i]) == Q.QDT_OK
                got[i] = O.canon_hash(t.triangles())
        except Exception as e:  # reported below
            errs.append(repr(e))

    th = [threading.Thread(target=work, args=(i,)) for i in range(8)]
    for x in th:
        x.start()
    for x in th:
        x.join()
    assert not errs, errs
    assert got == want


def test_context_reuse_and_growth(gpu):
    with Q.Triangulation() as t:
        for n in (10, 100000, 7, 300000, 3):
            p = W.uniform(n, n)
            print("reuse n=%d" % n, flush=True)
            assert t.execute(p) == Q.QDT_OK
            assert O.canon_hash(t.triangles()) == O.canon_hash(O.triangulate(p).tris)
