"""libqdt_amd -- MI355X (gfx950) batch Delaunay triangulation behind libqdt's C API.

The product is the shared library libqdt_amd/_lib/libqdt_amd.so, which
exports the fourteen functions of include/qdt.h (libqdt's public interface)
plus the GPU extensions of include/qdt_gpu.h.  This module loads it with
ctypes and mirrors the reference's context interface (libqdt/qdt.h) for
Python callers, tests and the bench.  There is no CPU fallback: without the
built library, or without a GPU, triangulation calls fail loudly.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libqdt_amd.so")

# qdt_status values (include/qdt.h, same as libqdt/qdt.h:79-93)
QDT_OK = 0
QDT_ERR_NULL_CONTEXT = -1
QDT_ERR_NULL_ARGUMENT = -2
QDT_ERR_TOO_MANY_POINTS = -3
QDT_ERR_NOT_FINITE = -4
QDT_ERR_INVALID_INDEX = -5
QDT_ERR_SAME_POINT = -6
QDT_ERR_NO_TRIANGLES = -7
QDT_ERR_CROSSES_CONSTRAINT = -8
QDT_ERR_THROUGH_POINT = -9
QDT_ERR_OUT_OF_MEMORY = -10
QDT_ERR_INTERNAL = -11

EXPORTS = [
    "qdt_status_string", "qdt_create", "qdt_execute", "qdt_execute_parallel", "qdt_insert_point",
    "qdt_insert_point_parallel", "qdt_insert_constraint", "qdt_insert_constraints_batch",
    "qdt_triangles_size", "qdt_get_triangle", "qdt_points_size", "qdt_get_point", "qdt_point_is_used",
    "qdt_destroy",
]
GPU_EXPORTS = [
    "qdt_gpu_execute_device", "qdt_gpu_device_triangles", "qdt_gpu_copy_triangles", "qdt_gpu_get_stats",
    "qdt_gpu_set_profiling", "qdt_gpu_set_seams", "qdt_gpu_set_cells", "qdt_gpu_set_test_limits",
    "qdt_gpu_execute_noalloc",
    "qdt_gpu_device_count",
    "qdt_gpu_release_workspace", "qdt_gpu_verify", "qdt_gpu_verify_triangles", "qdt_gpu_build_mesh",
    "qdt_gpu_batch_create", "qdt_gpu_batch_destroy", "qdt_gpu_batch_execute", "qdt_gpu_batch_execute_device",
    "qdt_gpu_batch_sets", "qdt_gpu_batch_set_status", "qdt_gpu_batch_triangles_size",
    "qdt_gpu_batch_unique_points", "qdt_gpu_batch_get_triangle", "qdt_gpu_batch_set_range",
    "qdt_gpu_batch_copy_triangles", "qdt_gpu_batch_device_triangles", "qdt_gpu_batch_copy_rep",
    "qdt_gpu_batch_get_stats", "qdt_gpu_batch_set_profiling", "qdt_gpu_batch_set_seams",
    "qdt_gpu_batch_set_graphs", "qdt_gpu_batch_set_tuning",
    "qdt_gpu_server_create", "qdt_gpu_server_destroy", "qdt_gpu_server_submit", "qdt_gpu_server_poll",
    "qdt_gpu_server_wait", "qdt_gpu_server_times", "qdt_gpu_server_triangles_size",
    "qdt_gpu_server_unique_points", "qdt_gpu_server_copy_triangles", "qdt_gpu_server_release",
    "qdt_gpu_server_counters", "qdt_gpu_server_memory",
]


class Vec2(ctypes.Structure):
    _fields_ = [("x", ctypes.c_double), ("y", ctypes.c_double)]


class Edge(ctypes.Structure):
    _fields_ = [("v0", ctypes.c_size_t), ("v1", ctypes.c_size_t)]


class GpuStats(ctypes.Structure):
    _fields_ = [
        ("orient_calls", ctypes.c_uint64), ("orient_exact", ctypes.c_uint64),
        ("incircle_calls", ctypes.c_uint64), ("incircle_exact", ctypes.c_uint64),
        ("extract_exact", ctypes.c_uint64), ("unique_points", ctypes.c_uint64),
        ("triangles", ctypes.c_uint64), ("kd_levels", ctypes.c_uint64), ("device_ms", ctypes.c_double),
        ("stage_ms", ctypes.c_double * 8), ("merge_launch_ms_sum", ctypes.c_double),
        ("merge_launches", ctypes.c_uint64), ("merge_edges", ctypes.c_uint64),
        ("seam_level_ms_sum", ctypes.c_double), ("seam_retries", ctypes.c_uint64),
        ("seam_counters", ctypes.c_uint64 * 8), ("err_bits", ctypes.c_uint64), ("cells", ctypes.c_uint64),
        ("cell_level", ctypes.c_uint64), ("spare_records", ctypes.c_uint64),
        ("cells_exact", ctypes.c_uint64), ("graph_replay", ctypes.c_uint64), ("graph_captured", ctypes.c_uint64),
        ("graph_fallback", ctypes.c_uint64), ("cell_points", ctypes.c_uint64), ("merge_lanes", ctypes.c_uint64),
        ("tune_source", ctypes.c_uint64), ("tune_ms", ctypes.c_double),
    ]
    STAGES = ("keys_sort_y", "sort_x", "dedupe", "frame1_sort", "kd_levels", "leaves", "merges", "extract")

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k not in ("stage_ms", "seam_counters")}
        d["stage_ms"] = dict(zip(self.STAGES, list(self.stage_ms)))
        d["seam"] = dict(zip(("certified", "inserted", "pieces", "resumed", "stopped", "dry_run_ok", "selected",
                                  "rings_rebuilt"),
                             list(self.seam_counters)))
        d["err_bits"] = int(self.err_bits)
        return d


_lib = None


class VerifyReport(ctypes.Structure):
    _fields_ = [
        ("triangles", ctypes.c_uint64), ("boundary_edges", ctypes.c_uint64), ("vertices", ctypes.c_uint64),
        ("bad_index", ctypes.c_uint64), ("not_representative", ctypes.c_uint64),
        ("bad_orientation", ctypes.c_uint64), ("bad_edges", ctypes.c_uint64), ("non_delaunay", ctypes.c_uint64),
        ("concave_boundary", ctypes.c_uint64), ("unused_points", ctypes.c_uint64), ("off_line", ctypes.c_uint64),
        ("euler_ok", ctypes.c_int),
        ("ok", ctypes.c_int), ("device_ms", ctypes.c_double),
    ]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["ok"] = bool(self.ok)
        d["euler_ok"] = bool(self.euler_ok)
        return d


class MeshInfo(ctypes.Structure):
    _fields_ = [
        ("vertices", ctypes.c_uint64), ("halfedges", ctypes.c_uint64), ("boundary", ctypes.c_uint64),
        ("unmatched", ctypes.c_uint64), ("device_ms", ctypes.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def lib():
    """The loaded C library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libqdt_amd.so is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        vp, sz, st = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int
        L.qdt_status_string.restype = ctypes.c_char_p
        L.qdt_status_string.argtypes = [st]
        L.qdt_create.restype = vp
        L.qdt_create.argtypes = []
        L.qdt_destroy.restype = None
        L.qdt_destroy.argtypes = [vp]
        for f in ("qdt_execute", "qdt_execute_parallel"):
            getattr(L, f).restype = st
            getattr(L, f).argtypes = [vp, vp, sz]
        for f in ("qdt_insert_point", "qdt_insert_point_parallel"):
            getattr(L, f).restype = st
            getattr(L, f).argtypes = [vp, Vec2, ctypes.POINTER(sz)]
        L.qdt_insert_constraint.restype = st
        L.qdt_insert_constraint.argtypes = [vp, sz, sz]
        L.qdt_insert_constraints_batch.restype = st
        L.qdt_insert_constraints_batch.argtypes = [vp, vp, sz, vp]
        L.qdt_triangles_size.restype = sz
        L.qdt_triangles_size.argtypes = [vp]
        L.qdt_get_triangle.restype = sz
        L.qdt_get_triangle.argtypes = [vp, ctypes.POINTER(sz), sz]
        L.qdt_points_size.restype = sz
        L.qdt_points_size.argtypes = [vp]
        L.qdt_get_point.restype = sz
        L.qdt_get_point.argtypes = [vp, ctypes.POINTER(Vec2), sz]
        L.qdt_point_is_used.restype = sz
        L.qdt_point_is_used.argtypes = [vp, sz]
        L.qdt_gpu_execute_device.restype = st
        L.qdt_gpu_execute_device.argtypes = [vp, vp, sz]
        L.qdt_gpu_device_triangles.restype = sz
        L.qdt_gpu_device_triangles.argtypes = [vp, ctypes.POINTER(vp)]
        L.qdt_gpu_copy_triangles.restype = sz
        L.qdt_gpu_copy_triangles.argtypes = [vp, vp]
        L.qdt_gpu_get_stats.restype = ctypes.c_int
        L.qdt_gpu_get_stats.argtypes = [vp, ctypes.POINTER(GpuStats)]
        L.qdt_gpu_set_profiling.restype = None
        L.qdt_gpu_set_profiling.argtypes = [vp, ctypes.c_int]
        L.qdt_gpu_set_seams.restype = None
        L.qdt_gpu_set_seams.argtypes = [vp, ctypes.c_int]
        L.qdt_gpu_set_cells.restype = None
        L.qdt_gpu_set_cells.argtypes = [vp, ctypes.c_int]
        L.qdt_gpu_set_test_limits.restype = None
        L.qdt_gpu_set_test_limits.argtypes = [vp, sz, sz]
        L.qdt_gpu_execute_noalloc.restype = st
        L.qdt_gpu_execute_noalloc.argtypes = [vp, sz, vp, sz, vp, ctypes.POINTER(sz)]
        L.qdt_gpu_verify.restype = st
        L.qdt_gpu_verify.argtypes = [vp, ctypes.POINTER(VerifyReport)]
        L.qdt_gpu_verify_triangles.restype = st
        L.qdt_gpu_verify_triangles.argtypes = [vp, sz, vp, sz, vp, ctypes.POINTER(VerifyReport)]
        L.qdt_gpu_build_mesh.restype = st
        L.qdt_gpu_build_mesh.argtypes = [vp, sz, sz, vp, vp, vp, ctypes.POINTER(MeshInfo)]
        L.qdt_gpu_device_count.restype = ctypes.c_int
        L.qdt_gpu_device_count.argtypes = []
        L.qdt_gpu_release_workspace.restype = None
        L.qdt_gpu_release_workspace.argtypes = [vp]
        # batched form
        L.qdt_gpu_batch_create.restype = vp
        L.qdt_gpu_batch_create.argtypes = []
        L.qdt_gpu_batch_destroy.restype = None
        L.qdt_gpu_batch_destroy.argtypes = [vp]
        for f in ("qdt_gpu_batch_execute", "qdt_gpu_batch_execute_device"):
            getattr(L, f).restype = st
            getattr(L, f).argtypes = [vp, vp, ctypes.POINTER(sz), sz]
        L.qdt_gpu_batch_sets.restype = sz
        L.qdt_gpu_batch_sets.argtypes = [vp]
        L.qdt_gpu_batch_set_status.restype = st
        L.qdt_gpu_batch_set_status.argtypes = [vp, sz]
        for f in ("qdt_gpu_batch_triangles_size", "qdt_gpu_batch_unique_points"):
            getattr(L, f).restype = sz
            getattr(L, f).argtypes = [vp, sz]
        L.qdt_gpu_batch_get_triangle.restype = sz
        L.qdt_gpu_batch_get_triangle.argtypes = [vp, sz, ctypes.POINTER(sz), sz]
        L.qdt_gpu_batch_set_range.restype = ctypes.c_int
        L.qdt_gpu_batch_set_range.argtypes = [vp, sz, ctypes.POINTER(sz), ctypes.POINTER(sz)]
        for f in ("qdt_gpu_batch_copy_triangles", "qdt_gpu_batch_copy_rep"):
            getattr(L, f).restype = sz
            getattr(L, f).argtypes = [vp, vp]
        L.qdt_gpu_batch_device_triangles.restype = sz
        L.qdt_gpu_batch_device_triangles.argtypes = [vp, ctypes.POINTER(vp)]
        L.qdt_gpu_batch_get_stats.restype = ctypes.c_int
        L.qdt_gpu_batch_get_stats.argtypes = [vp, ctypes.POINTER(GpuStats)]
        L.qdt_gpu_batch_set_tuning.restype = None
        L.qdt_gpu_batch_set_tuning.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        for f in ("qdt_gpu_batch_set_profiling", "qdt_gpu_batch_set_seams", "qdt_gpu_batch_set_graphs"):
            getattr(L, f).restype = None
            getattr(L, f).argtypes = [vp, ctypes.c_int]
        # serving form
        u64, dp = ctypes.c_uint64, ctypes.POINTER(ctypes.c_double)
        L.qdt_gpu_server_create.restype = vp
        L.qdt_gpu_server_create.argtypes = [sz, ctypes.c_int]
        L.qdt_gpu_server_destroy.restype = None
        L.qdt_gpu_server_destroy.argtypes = [vp]
        L.qdt_gpu_server_submit.restype = u64
        L.qdt_gpu_server_submit.argtypes = [vp, vp, sz]
        L.qdt_gpu_server_poll.restype = ctypes.c_int
        L.qdt_gpu_server_poll.argtypes = [vp, u64]
        L.qdt_gpu_server_wait.restype = st
        L.qdt_gpu_server_wait.argtypes = [vp, u64, dp]
        L.qdt_gpu_server_times.restype = ctypes.c_int
        L.qdt_gpu_server_times.argtypes = [vp, u64, dp, dp, dp]
        for f in ("qdt_gpu_server_triangles_size", "qdt_gpu_server_unique_points"):
            getattr(L, f).restype = sz
            getattr(L, f).argtypes = [vp, u64]
        L.qdt_gpu_server_copy_triangles.restype = sz
        L.qdt_gpu_server_copy_triangles.argtypes = [vp, u64, vp]
        L.qdt_gpu_server_release.restype = None
        L.qdt_gpu_server_release.argtypes = [vp, u64]
        L.qdt_gpu_server_counters.restype = ctypes.c_int
        L.qdt_gpu_server_counters.argtypes = [vp, ctypes.POINTER(u64), ctypes.POINTER(u64)]
        L.qdt_gpu_server_memory.restype = ctypes.c_int
        L.qdt_gpu_server_memory.argtypes = [vp, ctypes.POINTER(sz), ctypes.POINTER(sz)]
        _lib = L
    return _lib


def status_string(status):
    return lib().qdt_status_string(int(status)).decode()


def device_count():
    return int(lib().qdt_gpu_device_count())


def _as_points(pts):
    a = np.ascontiguousarray(pts, dtype=np.float64)
    if a.size == 0:
        return a.reshape(0, 2)
    return a.reshape(-1, 2)


class Triangulation:
    """One qdt_t context (libqdt/qdt.h:125): batch execute, then queries."""

    def __init__(self):
        self._h = lib().qdt_create()
        if not self._h:
            raise MemoryError("qdt_create returned NULL")

    def close(self):
        if self._h:
            lib().qdt_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def handle(self):
        return self._h

    # --- batch (libqdt/qdt.h:160,181) -------------------------------------
    def execute(self, pts, parallel=False):
        a = _as_points(pts)
        f = lib().qdt_execute_parallel if parallel else lib().qdt_execute
        return int(f(self._h, a.ctypes.data if len(a) else None, len(a)))

    def execute_parallel(self, pts):
        return self.execute(pts, parallel=True)

    def execute_device(self, dev_ptr, n):
        """Points already in GPU memory (e.g. a torch tensor's data_ptr())."""
        return int(lib().qdt_gpu_execute_device(self._h, ctypes.c_void_p(int(dev_ptr)), int(n)))

    # --- incremental (libqdt/qdt.h:219-314) ----------------------------------
    def insert_point(self, x, y, parallel=False):
        idx = ctypes.c_size_t(0)
        f = lib().qdt_insert_point_parallel if parallel else lib().qdt_insert_point
        st = int(f(self._h, Vec2(x, y), ctypes.byref(idx)))
        return st, idx.value

    def insert_constraint(self, v0, v1):
        return int(lib().qdt_insert_constraint(self._h, v0, v1))

    # --- queries (libqdt/qdt.h:326-394) --------------------------------------
    def triangles_size(self):
        return int(lib().qdt_triangles_size(self._h))

    def get_triangle(self, i):
        T = (ctypes.c_size_t * 3)()
        ok = lib().qdt_get_triangle(self._h, T, i)
        return (T[0], T[1], T[2]) if ok else None

    def points_size(self):
        return int(lib().qdt_points_size(self._h))

    def get_point(self, i):
        p = Vec2()
        ok = lib().qdt_get_point(self._h, ctypes.byref(p), i)
        return (p.x, p.y) if ok else None

    def point_is_used(self, i):
        return int(lib().qdt_point_is_used(self._h, i))

    # --- bulk / GPU extras (include/qdt_gpu.h) -------------------------------
    def triangles(self):
        """All triangles as a (T, 3) uint32 array (counter-clockwise)."""
        t = self.triangles_size()
        out = np.zeros((t, 3), dtype=np.uint32)
        if t:
            got = lib().qdt_gpu_copy_triangles(self._h, out.ctypes.data)
            assert got == t
        return out

    def device_triangles(self):
        """(device pointer, T) of the last batch's triangles."""
        p = ctypes.c_void_p()
        t = lib().qdt_gpu_device_triangles(self._h, ctypes.byref(p))
        return (p.value or 0), int(t)

    def stats(self):
        s = GpuStats()
        lib().qdt_gpu_get_stats(self._h, ctypes.byref(s))
        return s.as_dict()

    def verify(self):
        """Exact device check of the last result (qdt_gpu_verify): a dict of
        counts with 'ok'; raises on a failed call."""
        r = VerifyReport()
        st = lib().qdt_gpu_verify(self._h, ctypes.byref(r))
        if st != QDT_OK:
            raise RuntimeError("qdt_gpu_verify: " + status_string(st))
        return r.as_dict()

    def mesh(self):
        """Half-edge mesh of the last result, built on the device
        (qdt_gpu_build_mesh): (he (3T, 4) int32 origin/twin/next/triangle,
        vert (V, 2) int32 input index/half-edge, orig_to_local (n,) int32,
        info dict).  Device buffers come from torch."""
        import torch
        d_tris, T = self.device_triangles()
        n = self.points_size()
        he = torch.empty((max(3 * T, 1), 4), dtype=torch.int32, device="cuda")
        vert = torch.empty((max(min(n, 3 * T), 1), 2), dtype=torch.int32, device="cuda")
        o2l = torch.empty(max(n, 1), dtype=torch.int32, device="cuda")
        info = build_mesh(d_tris, T, n, he.data_ptr(), vert.data_ptr(), o2l.data_ptr())
        torch.cuda.synchronize()
        return (he[:3 * T].cpu().numpy(), vert[:info["vertices"]].cpu().numpy(), o2l[:n].cpu().numpy(), info)

    def set_profiling(self, on=True):
        lib().qdt_gpu_set_profiling(self._h, 1 if on else 0)

    def set_seams(self, on=True):
        lib().qdt_gpu_set_seams(self._h, 1 if on else 0)

    def set_cells(self, on=True):
        lib().qdt_gpu_set_cells(self._h, 1 if on else 0)

    def set_test_limits(self, device_bytes=0, spare_records=0):
        lib().qdt_gpu_set_test_limits(self._h, device_bytes, spare_records)

    def release_workspace(self):
        lib().qdt_gpu_release_workspace(self._h)


class Batch:
    """Many independent point sets in one GPU pass (include/qdt_gpu.h,
    qdt_gpu_batch_*).  Set i is points [offsets[i], offsets[i+1])."""

    def __init__(self):
        self._h = lib().qdt_gpu_batch_create()
        if not self._h:
            raise MemoryError("qdt_gpu_batch_create returned NULL")
        self._k = 0

    def close(self):
        if self._h:
            lib().qdt_gpu_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @staticmethod
    def _offsets(offsets):
        o = np.ascontiguousarray(offsets, dtype=np.uint64)
        return o, o.ctypes.data_as(ctypes.POINTER(ctypes.c_size_t)), max(len(o) - 1, 0)

    def execute(self, pts, offsets):
        """Host points (n, 2) and k + 1 offsets."""
        a = _as_points(pts)
        o, op, k = self._offsets(offsets)
        st = int(lib().qdt_gpu_batch_execute(self._h, a.ctypes.data if len(a) else None, op, k))
        self._k = k if st == QDT_OK else 0
        return st

    def execute_device(self, dev_ptr, offsets):
        """Points already in GPU memory; offsets on the host."""
        o, op, k = self._offsets(offsets)
        st = int(lib().qdt_gpu_batch_execute_device(self._h, ctypes.c_void_p(int(dev_ptr)), op, k))
        self._k = k if st == QDT_OK else 0
        return st

    def sets(self):
        return int(lib().qdt_gpu_batch_sets(self._h))

    def status(self, i):
        return int(lib().qdt_gpu_batch_set_status(self._h, i))

    def triangles_size(self, i):
        return int(lib().qdt_gpu_batch_triangles_size(self._h, i))

    def unique_points(self, i):
        return int(lib().qdt_gpu_batch_unique_points(self._h, i))

    def get_triangle(self, i, t):
        T = (ctypes.c_size_t * 3)()
        ok = lib().qdt_gpu_batch_get_triangle(self._h, i, T, t)
        return (T[0], T[1], T[2]) if ok else None

    def set_range(self, i):
        a, c = ctypes.c_size_t(), ctypes.c_size_t()
        ok = lib().qdt_gpu_batch_set_range(self._h, i, ctypes.byref(a), ctypes.byref(c))
        return (a.value, c.value) if ok else None

    def all_triangles(self):
        """(total, 3) uint32: every set's triangles at its set_range()."""
        p = ctypes.c_void_p()
        total = int(lib().qdt_gpu_batch_device_triangles(self._h, ctypes.byref(p)))
        out = np.zeros((total, 3), dtype=np.uint32)
        if total:
            got = lib().qdt_gpu_batch_copy_triangles(self._h, out.ctypes.data)
            assert got == total
        return out

    def triangles(self, i, all_tris=None):
        """(T, 3) uint32 triangles of set i, set-local indices."""
        r = self.set_range(i)
        if r is None:
            return None
        a = self.all_triangles() if all_tris is None else all_tris
        return a[r[0]:r[0] + r[1]]

    def rep(self, n):
        out = np.zeros(n, dtype=np.uint32)
        if n:
            lib().qdt_gpu_batch_copy_rep(self._h, out.ctypes.data)
        return out

    def stats(self):
        s = GpuStats()
        lib().qdt_gpu_batch_get_stats(self._h, ctypes.byref(s))
        return s.as_dict()

    def set_profiling(self, on=True):
        lib().qdt_gpu_batch_set_profiling(self._h, 1 if on else 0)

    def set_seams(self, on=True):
        lib().qdt_gpu_batch_set_seams(self._h, 1 if on else 0)

    def set_graphs(self, on=True):
        lib().qdt_gpu_batch_set_graphs(self._h, 1 if on else 0)

    def set_tuning(self, cell_points=0, merge_lanes=0):
        """0, 0: autotuned per shape bucket."""
        lib().qdt_gpu_batch_set_tuning(self._h, int(cell_points), int(merge_lanes))


class Server:
    """Request loop with continuous batching (include/qdt_gpu.h,
    qdt_gpu_server_*).  submit() may be called from any thread."""

    def __init__(self, hbm_budget_bytes=200 << 30, workers=2):
        self._h = lib().qdt_gpu_server_create(int(hbm_budget_bytes), int(workers))
        if not self._h:
            raise RuntimeError("qdt_gpu_server_create failed (no GPU?)")

    def close(self):
        if self._h:
            lib().qdt_gpu_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, pts):
        a = _as_points(pts)
        rid = int(lib().qdt_gpu_server_submit(self._h, a.ctypes.data if len(a) else None, len(a)))
        if not rid:
            raise MemoryError("qdt_gpu_server_submit failed")
        return rid

    def poll(self, rid):
        return int(lib().qdt_gpu_server_poll(self._h, rid))

    def wait(self, rid):
        """(status, latency in ms)"""
        lat = ctypes.c_double()
        st = int(lib().qdt_gpu_server_wait(self._h, rid, ctypes.byref(lat)))
        return st, lat.value

    def times(self, rid):
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        ok = lib().qdt_gpu_server_times(self._h, rid, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        return (a.value, b.value, c.value) if ok else None

    def triangles(self, rid):
        t = int(lib().qdt_gpu_server_triangles_size(self._h, rid))
        out = np.zeros((t, 3), dtype=np.uint32)
        if t:
            lib().qdt_gpu_server_copy_triangles(self._h, rid, out.ctypes.data)
        return out

    def unique_points(self, rid):
        return int(lib().qdt_gpu_server_unique_points(self._h, rid))

    def release(self, rid):
        lib().qdt_gpu_server_release(self._h, rid)

    def counters(self):
        b, r = ctypes.c_uint64(), ctypes.c_uint64()
        lib().qdt_gpu_server_counters(self._h, ctypes.byref(b), ctypes.byref(r))
        return b.value, r.value

    def memory(self):
        """(device bytes the workers hold now, largest sum so far)."""
        h, p = ctypes.c_size_t(), ctypes.c_size_t()
        lib().qdt_gpu_server_memory(self._h, ctypes.byref(h), ctypes.byref(p))
        return h.value, p.value


def execute_noalloc(d_pts, n, d_tris, tri_capacity, d_rep):
    """qdt_gpu_execute_noalloc on device pointers (integers, e.g. torch
    data_ptr()): (status, triangle count)."""
    t = ctypes.c_size_t(0)
    st = int(lib().qdt_gpu_execute_noalloc(ctypes.c_void_p(int(d_pts)), int(n), ctypes.c_void_p(int(d_tris)),
                                            int(tri_capacity), ctypes.c_void_p(int(d_rep)), ctypes.byref(t)))
    return st, t.value


def verify_triangles(d_pts, n, d_tris, T, d_rep=None):
    """Exact device check of T triangles (device uint32 array, 3 T input
    indices) over n device points (qdt_gpu_verify_triangles); pointers as
    integers (e.g. torch data_ptr()).  Returns a dict of counts with 'ok'."""
    r = VerifyReport()
    st = lib().qdt_gpu_verify_triangles(d_pts, n, d_tris, T, d_rep, ctypes.byref(r))
    if st != QDT_OK:
        raise RuntimeError("qdt_gpu_verify_triangles: " + status_string(st))
    return r.as_dict()


def build_mesh(d_tris, T, n, d_he, d_vert, d_orig_to_local):
    """Half-edge mesh of T device triangles over n input indices
    (qdt_gpu_build_mesh) into caller-allocated device arrays (3T x 4, up to
    min(n, 3T) x 2 and n int32); pointers as integers.  Returns the counts."""
    r = MeshInfo()
    st = lib().qdt_gpu_build_mesh(d_tris, T, n, d_he, d_vert, d_orig_to_local, ctypes.byref(r))
    if st != QDT_OK:
        raise RuntimeError("qdt_gpu_build_mesh: " + status_string(st))
    return r.as_dict()


def triangulate(pts):
    """Convenience: (T, 3) uint32 triangles of pts on the GPU."""
    with Triangulation() as t:
        st = t.execute(pts)
        if st != QDT_OK:
            raise RuntimeError("qdt_execute failed: %s" % status_string(st))
        return t.triangles()
