"""The C ABI without a GPU: the library loads, exports every symbol its
headers declare, and follows libqdt's validation and NULL rules
(libqdt/qdt.cpp:21-113, 145-304) for calls that fail before any GPU work."""
import ctypes
import os
import re

import numpy as np
import pytest

import libqdt_amd as Q

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    return re.findall(r"QDT_API[^;]*?\b(qdt_\w+)\s*\(", txt)


def test_exports_every_declared_symbol():
    L = Q.lib()
    names = declared("qdt.h") + declared("qdt_gpu.h")
    assert len(declared("qdt.h")) == 14
    for n in names:
        assert hasattr(L, n), n
    assert sorted(declared("qdt.h")) == sorted(Q.EXPORTS)
    assert sorted(declared("qdt_gpu.h")) == sorted(Q.GPU_EXPORTS)


def test_status_strings():
    texts = set()
    for s in range(-11, 1):
        t = Q.status_string(s)
        assert t != "unknown status"
        texts.add(t)
    assert len(texts) == 12
    assert Q.status_string(12345) == "unknown status"
    assert Q.status_string(0) == "success"


def test_null_context():
    L = Q.lib()
    assert L.qdt_execute(None, None, 0) == Q.QDT_ERR_NULL_CONTEXT
    assert L.qdt_execute_parallel(None, None, 0) == Q.QDT_ERR_NULL_CONTEXT
    idx = ctypes.c_size_t(777)
    assert L.qdt_insert_point(None, Q.Vec2(0, 0), ctypes.byref(idx)) == Q.QDT_ERR_NULL_CONTEXT and idx.value == 777
    assert L.qdt_insert_constraint(None, 0, 1) == Q.QDT_ERR_NULL_CONTEXT
    res = (ctypes.c_int * 2)(0, 0)
    edges = (Q.Edge * 2)(Q.Edge(0, 1), Q.Edge(1, 2))
    assert L.qdt_insert_constraints_batch(None, edges, 2, res) == Q.QDT_ERR_NULL_CONTEXT
    assert list(res) == [Q.QDT_ERR_NULL_CONTEXT] * 2
    assert L.qdt_triangles_size(None) == 0 and L.qdt_points_size(None) == 0
    T = (ctypes.c_size_t * 3)(7, 8, 9)
    assert L.qdt_get_triangle(None, T, 0) == 0 and list(T) == [7, 8, 9]
    p = Q.Vec2(5, 6)
    assert L.qdt_get_point(None, ctypes.byref(p), 0) == 0 and (p.x, p.y) == (5, 6)
    assert L.qdt_point_is_used(None, 0) == 0
    L.qdt_destroy(None)


def test_fresh_context_and_rejections_need_no_gpu():
    with Q.Triangulation() as t:
        assert t.triangles_size() == 0 and t.points_size() == 0
        ok = np.array([[0, 0], [4, 0], [3, 1], [1, 1]], np.float64)
        L = Q.lib()
        # TOO_MANY_POINTS is checked before the array is read, even a NULL one
        assert L.qdt_execute(t.handle, ok.ctypes.data, 100000001) == Q.QDT_ERR_TOO_MANY_POINTS
        assert L.qdt_execute(t.handle, None, 100000001) == Q.QDT_ERR_TOO_MANY_POINTS
        assert L.qdt_execute(t.handle, None, 4) == Q.QDT_ERR_NULL_ARGUMENT
        for bad in ([np.nan, 1], [np.inf, 1], [1, -np.inf]):
            p = ok.copy()
            p[2] = bad
            assert t.execute(p) == Q.QDT_ERR_NOT_FINITE
            assert t.execute_parallel(p) == Q.QDT_ERR_NOT_FINITE
            assert t.points_size() == 0 and t.triangles_size() == 0
        assert t.execute(np.zeros((0, 2))) == Q.QDT_OK
        assert t.points_size() == 0
        # rejected single points leave the context as it was
        assert t.insert_point(float("nan"), 0.0)[0] == Q.QDT_ERR_NOT_FINITE
        assert t.insert_point(0.0, float("inf"))[0] == Q.QDT_ERR_NOT_FINITE
        assert t.insert_constraint(0, 1) == Q.QDT_ERR_INVALID_INDEX
        assert t.get_triangle(0) is None and t.get_point(0) is None
        assert t.point_is_used(0) == 0
        L.qdt_insert_constraints_batch(t.handle, None, 0, None)
        res = (ctypes.c_int * 3)(0, 0, 0)
        assert L.qdt_insert_constraints_batch(t.handle, None, 3, res) == Q.QDT_ERR_NULL_ARGUMENT
        assert list(res) == [Q.QDT_ERR_NULL_ARGUMENT] * 3


def test_no_gpu_is_a_loud_error():
    if Q.device_count() > 0:
        pytest.skip("a GPU is present")
    with Q.Triangulation() as t:
        st = t.execute(np.array([[0, 0], [1, 0], [0, 1]], np.float64))
        assert st == Q.QDT_ERR_INTERNAL
        assert t.points_size() == 0 and t.triangles_size() == 0


def test_batch_validation_needs_no_gpu():
    """qdt_gpu_batch_*: offsets and pointer checks come before any device
    work; with no GPU the call itself fails loudly (no CPU fallback)."""
    import numpy as np

    b = Q.Batch()
    pts = np.zeros((4, 2))
    assert b.execute(pts, [1, 4]) == Q.QDT_ERR_INVALID_INDEX
    assert b.execute(pts, [0, 3, 2]) == Q.QDT_ERR_INVALID_INDEX
    assert b.execute(pts, [0]) == Q.QDT_OK and b.sets() == 0
    L = Q.lib()
    assert L.qdt_gpu_batch_execute(None, None, None, 0) == Q.QDT_ERR_NULL_CONTEXT
    assert L.qdt_gpu_batch_execute(b._h, None, None, 2) == Q.QDT_ERR_NULL_ARGUMENT
    assert b.status(0) == Q.QDT_ERR_INVALID_INDEX
    assert b.triangles_size(0) == 0 and b.get_triangle(0, 0) is None
    if not Q.device_count():
        assert b.execute(pts, [0, 4]) == Q.QDT_ERR_INTERNAL
        assert b.sets() == 0
    b.close()


def test_server_needs_a_gpu():
    if Q.device_count():
        return
    assert Q.lib().qdt_gpu_server_create(1 << 30, 2) is None
    assert Q.lib().qdt_gpu_server_poll(None, 1) == -1
    assert Q.lib().qdt_gpu_server_wait(None, 1, None) == Q.QDT_ERR_NULL_CONTEXT


def test_mesh_argument_checks_need_no_gpu():
    """qdt_gpu_build_mesh: NULL and size checks come before any device work."""
    import ctypes

    L = Q.lib()
    info = Q.MeshInfo()
    assert L.qdt_gpu_build_mesh(None, 0, 0, None, None, None, None) == Q.QDT_ERR_NULL_ARGUMENT
    assert L.qdt_gpu_build_mesh(None, 5, 10, None, None, None, ctypes.byref(info)) == Q.QDT_ERR_NULL_ARGUMENT
    assert L.qdt_gpu_build_mesh(None, 0, 10, None, None, None, ctypes.byref(info)) == Q.QDT_ERR_NULL_ARGUMENT
    big = 500_000_000
    assert L.qdt_gpu_build_mesh(8, big, 10, 8, 8, 8, ctypes.byref(info)) == Q.QDT_ERR_TOO_MANY_POINTS
