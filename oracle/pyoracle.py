"""ctypes access to the CPU oracle (liboracle.so) and to the compiled
reference checker (oracle/_ref/libqdt_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
bench's CPU baseline leg, never by the product package.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_ORACLE = os.path.join(HERE, "liboracle.so")
_REF = os.path.join(HERE, "_ref", "libqdt_ref.so")

_lib = None
_ref = None


def build(ref_root="/root/reference"):
    """Compile the oracle, and the reference checker when its sources exist."""
    subprocess.check_call(["make", "-s", "-C", HERE, "all"])
    if os.path.isdir(os.path.join(ref_root, "libqdt")):
        subprocess.check_call(["make", "-s", "-C", HERE, "ref", "REF=" + ref_root])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_ORACLE):
            build()
        L = ctypes.CDLL(_ORACLE)
        d = ctypes.c_double
        L.orc_orient2d.restype = d
        L.orc_orient2d.argtypes = [d] * 6
        L.orc_incircle.restype = d
        L.orc_incircle.argtypes = [d] * 8
        L.orc_orient2d_exact.restype = ctypes.c_int
        L.orc_orient2d_exact.argtypes = [d] * 6
        L.orc_incircle_exact.restype = ctypes.c_int
        L.orc_incircle_exact.argtypes = [d] * 8
        L.orc_triangulate.restype = ctypes.c_void_p
        L.orc_triangulate.argtypes = [ctypes.c_void_p, ctypes.c_size_t]
        L.orc_ntri.restype = ctypes.c_size_t
        L.orc_ntri.argtypes = [ctypes.c_void_p]
        L.orc_tris.restype = ctypes.POINTER(ctypes.c_uint32)
        L.orc_tris.argtypes = [ctypes.c_void_p]
        L.orc_rep.restype = ctypes.POINTER(ctypes.c_uint32)
        L.orc_rep.argtypes = [ctypes.c_void_p]
        L.orc_unique.restype = ctypes.c_size_t
        L.orc_unique.argtypes = [ctypes.c_void_p]
        L.orc_stats.restype = None
        L.orc_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64)]
        L.orc_free.restype = None
        L.orc_free.argtypes = [ctypes.c_void_p]
        L.orc_canon_hash.restype = ctypes.c_uint64
        L.orc_canon_hash.argtypes = [ctypes.c_void_p, ctypes.c_size_t]
        L.orc_canonicalize.restype = None
        L.orc_canonicalize.argtypes = [ctypes.c_void_p, ctypes.c_size_t]
        L.orc_hull.restype = ctypes.c_size_t
        L.orc_hull.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
        L.orc_check.restype = ctypes.c_int
        L.orc_check.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t,
                                ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_int64)]
        L.orc_build_mesh.restype = ctypes.c_long
        L.orc_build_mesh.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.c_void_p]
        _lib = L
    return _lib


def _xy(pts):
    a = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 2)
    return a


class Result:
    def __init__(self, tris, rep, unique, stats):
        self.tris = tris      # (T, 3) uint32 original indices, CCW
        self.rep = rep        # (n,) uint32
        self.unique = unique  # distinct points
        self.stats = stats    # orient calls, orient exact, incircle calls, incircle exact, records, live


def triangulate(pts):
    """The oracle's batch triangulation (restates qdt_execute)."""
    L = lib()
    a = _xy(pts)
    n = a.shape[0]
    h = L.orc_triangulate(a.ctypes.data if n else None, n)
    if not h:
        raise MemoryError("oracle out of memory")
    try:
        nt = L.orc_ntri(h)
        tris = np.ctypeslib.as_array(L.orc_tris(h), shape=(nt * 3,)).reshape(nt, 3).copy() if nt else \
            np.zeros((0, 3), np.uint32)
        rep = np.ctypeslib.as_array(L.orc_rep(h), shape=(n,)).copy() if n else np.zeros(0, np.uint32)
        st = (ctypes.c_uint64 * 6)()
        L.orc_stats(h, st)
        return Result(tris, rep, int(L.orc_unique(h)), list(st))
    finally:
        L.orc_free(h)


def canonical(tris):
    """Rotate each triangle to its smallest index and sort the list."""
    t = np.ascontiguousarray(tris, dtype=np.uint32).reshape(-1, 3).copy()
    if len(t):
        lib().orc_canonicalize(t.ctypes.data, len(t))
    return t


def canon_hash(tris):
    t = np.ascontiguousarray(tris, dtype=np.uint32).reshape(-1, 3)
    return int(lib().orc_canon_hash(t.ctypes.data if len(t) else None, len(t)))


def hull(pts):
    a = _xy(pts)
    out = np.zeros(2 * len(a) + 2, np.uint32)
    k = lib().orc_hull(a.ctypes.data if len(a) else None, len(a), out.ctypes.data)
    return out[:k].copy()


def check(pts, tris):
    """Exact validity check; returns (code, message, info) with code 0 = valid."""
    a = _xy(pts)
    t = np.ascontiguousarray(tris, dtype=np.uint32).reshape(-1, 3)
    msg = ctypes.create_string_buffer(256)
    info = (ctypes.c_int64 * 4)()
    rc = lib().orc_check(a.ctypes.data if len(a) else None, len(a), t.ctypes.data if len(t) else None,
                         len(t), msg, 256, info)
    return rc, msg.value.decode(), dict(m=info[0], b=info[1], T=info[2], E=info[3])


def orient2d(a, b, c):
    return lib().orc_orient2d(a[0], a[1], b[0], b[1], c[0], c[1])


def incircle(a, b, c, d):
    return lib().orc_incircle(a[0], a[1], b[0], b[1], c[0], c[1], d[0], d[1])


# --------------------------------------------------------------------------
# The compiled reference (built from /root/reference by oracle/Makefile).
# --------------------------------------------------------------------------

def ref_available():
    return os.path.exists(_REF)


class _Vec2(ctypes.Structure):
    _fields_ = [("x", ctypes.c_double), ("y", ctypes.c_double)]


def ref():
    global _ref
    if _ref is None:
        R = ctypes.CDLL(_REF)  # RTLD_LOCAL: its qdt_* names stay private
        vp = ctypes.c_void_p
        sz = ctypes.c_size_t
        R.qdt_create.restype = vp
        R.qdt_create.argtypes = []
        R.qdt_destroy.restype = None
        R.qdt_destroy.argtypes = [vp]
        R.qdt_execute.restype = ctypes.c_int
        R.qdt_execute.argtypes = [vp, vp, sz]
        R.qdt_execute_parallel.restype = ctypes.c_int
        R.qdt_execute_parallel.argtypes = [vp, vp, sz]
        R.qdt_triangles_size.restype = sz
        R.qdt_triangles_size.argtypes = [vp]
        R.qdt_get_triangle.restype = sz
        R.qdt_get_triangle.argtypes = [vp, ctypes.POINTER(sz), sz]
        R.qdt_points_size.restype = sz
        R.qdt_points_size.argtypes = [vp]
        R.qdt_point_is_used.restype = sz
        R.qdt_point_is_used.argtypes = [vp, sz]
        d = ctypes.c_double
        R.ref_orient2d.restype = d
        R.ref_orient2d.argtypes = [d] * 6
        R.ref_incircle.restype = d
        R.ref_incircle.argtypes = [d] * 8
        R.ref_copy_triangles.restype = sz
        R.ref_copy_triangles.argtypes = [vp, vp]
        R.qdt_insert_point.restype = ctypes.c_int
        R.qdt_insert_point.argtypes = [vp, _Vec2, ctypes.POINTER(sz)]
        R.qdt_insert_constraint.restype = ctypes.c_int
        R.qdt_insert_constraint.argtypes = [vp, sz, sz]
        _ref = R
    return _ref


def ref_triangulate(pts, parallel=False, timed=False):
    """Triangles (T, 3) from the compiled reference; optionally the call time in s."""
    import time
    R = ref()
    a = _xy(pts)
    n = a.shape[0]
    q = R.qdt_create()
    try:
        t0 = time.perf_counter()
        st = (R.qdt_execute_parallel if parallel else R.qdt_execute)(q, a.ctypes.data if n else None, n)
        dt = time.perf_counter() - t0
        if st != 0:
            raise RuntimeError("reference returned status %d" % st)
        nt = R.qdt_triangles_size(q)
        out = np.zeros((nt, 3), np.uint32)
        if nt:
            R.ref_copy_triangles(q, out.ctypes.data)
        used = np.array([R.qdt_point_is_used(q, i) for i in range(n)], np.uint8) if n <= 200000 else None
        return (out, used, dt) if timed else (out, used)
    finally:
        R.qdt_destroy(q)


class RefContext:
    """A context of the compiled reference, for the insertion and
    constraint parity tests: execute / insert / constrain / triangles."""

    def __init__(self):
        self.R = ref()
        self.q = self.R.qdt_create()

    def close(self):
        if self.q:
            self.R.qdt_destroy(self.q)
            self.q = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def execute(self, pts):
        a = _xy(pts)
        return self.R.qdt_execute(self.q, a.ctypes.data if len(a) else None, len(a))

    def insert(self, x, y):
        idx = ctypes.c_size_t(0)
        st = self.R.qdt_insert_point(self.q, _Vec2(float(x), float(y)), ctypes.byref(idx))
        return st, idx.value

    def constrain(self, a, b):
        return self.R.qdt_insert_constraint(self.q, int(a), int(b))

    def triangles(self):
        nt = self.R.qdt_triangles_size(self.q)
        out = np.zeros((nt, 3), np.uint32)
        if nt:
            self.R.ref_copy_triangles(self.q, out.ctypes.data)
        return out

    def used(self):
        n = self.R.qdt_points_size(self.q)
        return np.array([self.R.qdt_point_is_used(self.q, i) for i in range(n)], np.uint8)


def ref_batch_time(pts, offsets, threads):
    """Wall time (s) of the reference triangulating every set [offsets[i],
    offsets[i+1]) with qdt_execute on `threads` host threads (one context per
    thread), and the total triangle count."""
    R = ref()
    a = _xy(pts)
    o = np.ascontiguousarray(offsets, dtype=np.uint64)
    tot = ctypes.c_uint64()
    R.ref_batch_seconds.restype = ctypes.c_double
    R.ref_batch_seconds.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_uint64)]
    dt = R.ref_batch_seconds(a.ctypes.data, o.ctypes.data, len(o) - 1, int(threads), ctypes.byref(tot))
    if dt < 0:
        raise RuntimeError("reference failed on a set")
    return dt, tot.value


def ref_batch_hashes(pts, offsets, threads=8):
    """Per set [offsets[i], offsets[i+1]): the reference's triangle count and
    the FNV-1a-64 of its canonical triangle list (computed in the shim, a
    second implementation beside canon_hash)."""
    R = ref()
    a = _xy(pts)
    o = np.ascontiguousarray(offsets, dtype=np.uint64)
    k = len(o) - 1
    cnt = np.zeros(max(k, 1), np.uint64)
    hs = np.zeros(max(k, 1), np.uint64)
    R.ref_batch_hashes.restype = ctypes.c_int
    R.ref_batch_hashes.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int,
                                   ctypes.c_void_p, ctypes.c_void_p]
    if R.ref_batch_hashes(a.ctypes.data, o.ctypes.data, k, int(threads), cnt.ctypes.data, hs.ctypes.data) != 0:
        raise RuntimeError("reference failed on a set")
    return cnt[:k], hs[:k]


def ref_mesh_time(pts, parallel=True):
    """Wall time (s) of the reference's half-edge mesh build (ensure_mesh_ /
    ensure_mesh_parallel_) after an untimed batch call on pts, and 3 T."""
    R = ref()
    a = _xy(pts)
    he = ctypes.c_uint64()
    R.ref_mesh_seconds.restype = ctypes.c_double
    R.ref_mesh_seconds.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(ctypes.c_uint64)]
    dt = R.ref_mesh_seconds(a.ctypes.data, a.shape[0], 1 if parallel else 0, ctypes.byref(he))
    if dt < 0:
        raise RuntimeError("reference mesh build failed")
    return dt, he.value


def ref_time(pts, parallel=True):
    """Wall time (s) of one reference batch call, without reading results."""
    import time
    R = ref()
    a = _xy(pts)
    n = a.shape[0]
    q = R.qdt_create()
    try:
        t0 = time.perf_counter()
        st = (R.qdt_execute_parallel if parallel else R.qdt_execute)(q, a.ctypes.data if n else None, n)
        dt = time.perf_counter() - t0
        if st != 0:
            raise RuntimeError("reference returned status %d" % st)
        return dt, R.qdt_triangles_size(q)
    finally:
        R.qdt_destroy(q)


def build_mesh(tris, n):
    """The oracle's half-edge mesh of a triangle list (restates ensure_mesh_):
    (he (3T, 4) int32 origin/twin/next/triangle, vert (V, 2) int32 input
    index/half-edge, o2l (n,) int32)."""
    t = np.ascontiguousarray(tris, dtype=np.uint32).reshape(-1, 3)
    T = len(t)
    he = np.zeros((3 * T, 4), np.int32)
    vert = np.zeros((max(min(n, 3 * T), 1), 2), np.int32)
    o2l = np.zeros(max(n, 1), np.int32)
    nv = lib().orc_build_mesh(t.ctypes.data if T else None, T, n, he.ctypes.data, vert.ctypes.data, o2l.ctypes.data)
    if nv < 0:
        raise ValueError("orc_build_mesh: index out of range" if nv == -1 else "orc_build_mesh: out of memory")
    return he, vert[:nv].copy(), o2l[:n].copy()
