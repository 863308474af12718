// qdt_capi.cpp -- the qdt.h C ABI over the GPU engine.
//
// Contract (libqdt/qdt.cpp): input checks in the order TOO_MANY_POINTS ->
// NULL_ARGUMENT -> NOT_FINITE (qdt.cpp:29-43); a rejected, out-of-memory or
// internally failed batch leaves the context empty (qdt.cpp:56-59, 88-113);
// no C++ exception crosses the ABI; queries are O(1) reads of host memory
// (qdt.cpp:246-298).  The host store keeps triangles as uint32 (indices are
// below 1e8) and widens them on read.
#include "../../include/qdt.h"
#include "../../include/qdt_gpu.h"
#include "engine.hpp"
#include "mesh_edit.hpp"
#include "tuner.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <new>
#include <set>
#include <thread>
#include <utility>
#include <vector>

namespace
{

constexpr size_t kMaxPoints = 100000000; // libqdt/wrapper.hxx:15

// Growable pinned host buffer (fast DMA for H2D/D2H).
template <class T>
struct Pinned
{
    T* p = nullptr;
    size_t cap = 0;
    ~Pinned() { reset(); }
    void reset()
    {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
    }
    // grows to at least n elements, keeping the first `keep` ones
    bool grow_keep(size_t n, size_t keep)
    {
        if (n <= cap) return true;
        size_t c = std::max<size_t>(n, 2 * cap);
        void* q = nullptr;
        if (hipHostMalloc(&q, c * sizeof(T), hipHostMallocDefault) != hipSuccess) return false;
        if (keep) std::memcpy(q, p, keep * sizeof(T));
        reset();
        p = (T*)q;
        cap = c;
        return true;
    }
    bool reserve(size_t n)
    {
        if (n <= cap) return true;
        reset();
        size_t c = std::max<size_t>(n, 1024);
        void* q = nullptr;
        if (hipHostMalloc(&q, c * sizeof(T), hipHostMallocDefault) != hipSuccess) return false;
        p = (T*)q;
        cap = c;
        return true;
    }
};

} // namespace

struct qdt_t
{
    // host store
    Pinned<qdt_vec2d> pts;
    size_t npts = 0;
    Pinned<uint32_t> tris; // 3 per triangle
    size_t ntris = 0;
    Pinned<uint32_t> rep;
    size_t nrep = 0;
    // device side
    qdt::Engine* eng = nullptr;
    double* d_in = nullptr;
    size_t d_in_cap = 0;
    // lazy host copies after qdt_gpu_execute_device
    bool tris_on_device_only = false;
    bool pts_on_device_only = false;
    qdt_gpu_stats stats{};
    bool profiling = false;
    bool seams = true;
    bool cells = true;
    size_t mem_cap = 0, spare_limit = 0; // test limits (qdt_gpu_set_test_limits)
    // Insertions and constraints after a batch call work on a host mesh
    // that takes over the points, triangles and representatives (null until
    // the first such call on a triangulation with triangles).
    qdt::MeshEdit* ed = nullptr;

    ~qdt_t()
    {
        delete ed;
        delete eng;
        if (d_in) (void)hipFree(d_in);
    }
};

namespace
{

void reset_context(qdt_t* Q) noexcept
{
    Q->npts = 0;
    Q->ntris = 0;
    Q->nrep = 0;
    Q->tris_on_device_only = false;
    Q->pts_on_device_only = false;
    Q->stats = qdt_gpu_stats{};
    delete Q->ed;
    Q->ed = nullptr;
}

bool finite_pt(const qdt_vec2d& p) { return std::isfinite(p.x) && std::isfinite(p.y); }

// Runs f(begin, end) over [0, n) on up to 16 host threads (one for small n).
template <class F>
void parallel_ranges(size_t n, F f)
{
    const size_t kChunk = 1 << 20;
    unsigned nt = std::thread::hardware_concurrency();
    nt = std::min<unsigned>(std::max(1u, nt), 16u);
    if (n < 2 * kChunk || nt == 1)
    {
        f((size_t)0, n);
        return;
    }
    std::vector<std::thread> th;
    const size_t per = (n + nt - 1) / nt;
    size_t done = 0;
    try
    {
        for (unsigned t = 0; t < nt; ++t)
        {
            const size_t a = t * per, b = std::min(n, a + per);
            if (a >= b) break;
            th.emplace_back(f, a, b);
            done = b;
        }
    }
    catch (...)
    {
        // could not start a thread: do the rest here
    }
    for (auto& t : th) t.join();
    if (done < n) f(done, n);
}

// The reference scans every coordinate before doing anything else
// (qdt.cpp:37-41); this check allocates nothing.
bool all_finite(const qdt_vec2d* src, size_t n)
{
    std::atomic<bool> ok{ true };
    parallel_ranges(n, [&](size_t a, size_t b) {
        bool good = true;
        for (size_t i = a; i < b; ++i) good &= finite_pt(src[i]);
        if (!good) ok = false;
    });
    return ok;
}

void copy_points(const qdt_vec2d* src, size_t n, qdt_vec2d* dst)
{
    parallel_ranges(n, [&](size_t a, size_t b) { std::memcpy(dst + a, src + a, (b - a) * sizeof(qdt_vec2d)); });
}

qdt_status map_status(qdt::Status s)
{
    switch (s)
    {
    case qdt::Status::ok: return QDT_OK;
    case qdt::Status::not_finite: return QDT_ERR_NOT_FINITE;
    case qdt::Status::out_of_memory: return QDT_ERR_OUT_OF_MEMORY;
    default: return QDT_ERR_INTERNAL;
    }
}

qdt_status ensure_engine(qdt_t* Q)
{
    if (Q->eng) return QDT_OK;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return QDT_ERR_INTERNAL;
    Q->eng = new (std::nothrow) qdt::Engine();
    if (!Q->eng) return QDT_ERR_OUT_OF_MEMORY;
    Q->eng->set_profiling(Q->profiling);
    Q->eng->set_seams(Q->seams);
    Q->eng->set_cells(Q->cells);
    Q->eng->set_limits(Q->mem_cap, Q->spare_limit);
    if (!Q->eng->stream())
    {
        delete Q->eng;
        Q->eng = nullptr;
        return QDT_ERR_INTERNAL;
    }
    return QDT_OK;
}

qdt_status ensure_device_input(qdt_t* Q, size_t n)
{
    const size_t bytes = n * sizeof(qdt_vec2d);
    if (bytes <= Q->d_in_cap) return QDT_OK;
    if (Q->d_in) (void)hipFree(Q->d_in);
    Q->d_in = nullptr;
    Q->d_in_cap = 0;
    void* p = nullptr;
    const hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) return e == hipErrorOutOfMemory ? QDT_ERR_OUT_OF_MEMORY : QDT_ERR_INTERNAL;
    Q->d_in = (double*)p;
    Q->d_in_cap = bytes;
    return QDT_OK;
}

void copy_stats(const qdt::RunStats& s, qdt_gpu_stats& o)
{
    o.orient_calls = s.orient;
    o.orient_exact = s.orient_exact;
    o.incircle_calls = s.incircle;
    o.incircle_exact = s.incircle_exact;
    o.extract_exact = s.extract_queue;
    o.unique_points = s.unique;
    o.triangles = s.triangles;
    o.kd_levels = s.levels;
    o.device_ms = s.device_ms;
    for (int i = 0; i < 8; ++i) o.stage_ms[i] = s.stage_ms[i];
    o.merge_launch_ms_sum = s.merge_launch_ms_sum;
    o.merge_launches = s.merge_launches;
    o.merge_edges = s.merge_edges;
    o.seam_level_ms_sum = s.seam_level_ms_sum;
    o.seam_retries = s.seam_retries;
    for (int i = 0; i < 8; ++i) o.seam_counters[i] = s.seam[i];
    o.err_bits = s.err_bits;
    o.cells = s.cells;
    o.cell_level = s.block_level;
    o.spare_records = s.spare_records;
    o.cells_exact = s.cells_exact;
    o.graph_replay = s.graph_replay;
    o.graph_captured = s.graph_captured;
    o.graph_fallback = s.graph_fallback;
    o.cell_points = s.cell_points_max;
    o.merge_lanes = s.merge_lanes;
}

void record_stats(qdt_t* Q)
{
    copy_stats(Q->eng->stats(), Q->stats);
}

// Brings triangles and rep of the last device run into host memory.
qdt_status materialize(qdt_t* Q)
{
    if (!Q->tris_on_device_only) return QDT_OK;
    const qdt::DeviceResult& r = Q->eng->result();
    if (!Q->tris.reserve(3 * (size_t)r.triangles + 3) || !Q->rep.reserve(r.n + 1)) return QDT_ERR_OUT_OF_MEMORY;
    hipStream_t s = Q->eng->stream();
    if (r.triangles &&
        hipMemcpyAsync(Q->tris.p, r.tris, 12 * (size_t)r.triangles, hipMemcpyDeviceToHost, s) != hipSuccess)
        return QDT_ERR_INTERNAL;
    if (r.n && hipMemcpyAsync(Q->rep.p, r.rep, 4 * r.n, hipMemcpyDeviceToHost, s) != hipSuccess)
        return QDT_ERR_INTERNAL;
    if (hipStreamSynchronize(s) != hipSuccess) return QDT_ERR_INTERNAL;
    Q->ntris = r.triangles;
    Q->nrep = r.n;
    Q->tris_on_device_only = false;
    return QDT_OK;
}

qdt_status materialize_points(qdt_t* Q)
{
    if (!Q->pts_on_device_only) return QDT_OK;
    if (!Q->pts.reserve(Q->npts + 1)) return QDT_ERR_OUT_OF_MEMORY;
    if (Q->npts &&
        hipMemcpy(Q->pts.p, Q->d_in, Q->npts * sizeof(qdt_vec2d), hipMemcpyDeviceToHost) != hipSuccess)
        return QDT_ERR_INTERNAL;
    Q->pts_on_device_only = false;
    return QDT_OK;
}

// Runs the engine on the points already in Q->d_in (n of them).
qdt_status run_device(qdt_t* Q, size_t n)
{
    const qdt::Status s = Q->eng->run(Q->d_in, n);
    if (s != qdt::Status::ok) return map_status(s);
    record_stats(Q);
    Q->tris_on_device_only = true;
    return QDT_OK;
}

qdt_status batch_checks(const qdt_vec2d* pts, size_t n)
{
    if (n > kMaxPoints) return QDT_ERR_TOO_MANY_POINTS;
    if (n == 0) return QDT_OK;
    if (!pts) return QDT_ERR_NULL_ARGUMENT;
    return QDT_OK;
}

// qdt_execute / qdt_execute_parallel (libqdt/qdt.cpp:88-113)
qdt_status execute_host(qdt_t* Q, const qdt_vec2d* pts, size_t n)
{
    if (!Q) return QDT_ERR_NULL_CONTEXT;
    try
    {
        qdt_status st = batch_checks(pts, n);
        if (st != QDT_OK || n == 0)
        {
            reset_context(Q);
            return st;
        }
        reset_context(Q);
        if (!all_finite(pts, n)) return QDT_ERR_NOT_FINITE;
        st = ensure_engine(Q);
        if (st != QDT_OK) return st;
        if (!Q->pts.reserve(n)) throw std::bad_alloc();
        copy_points(pts, n, Q->pts.p);
        st = ensure_device_input(Q, n);
        if (st == QDT_OK &&
            hipMemcpyAsync(Q->d_in, Q->pts.p, n * sizeof(qdt_vec2d), hipMemcpyHostToDevice, Q->eng->stream()) !=
                hipSuccess)
            st = QDT_ERR_INTERNAL;
        if (st == QDT_OK) st = run_device(Q, n);
        if (st == QDT_OK) st = materialize(Q);
        if (st != QDT_OK)
        {
            reset_context(Q);
            return st;
        }
        Q->npts = n;
        return QDT_OK;
    }
    catch (const std::bad_alloc&)
    {
        reset_context(Q);
        return QDT_ERR_OUT_OF_MEMORY;
    }
    catch (...)
    {
        reset_context(Q);
        return QDT_ERR_INTERNAL;
    }
}

// Above this many triangles the half-edge twins for the host mesh are paired
// on the device (a sort of edge keys) instead of by the host.
constexpr size_t kDeviceTwinsMin = 1 << 16;

// The host mesh for insertions and constraints, built from the batch result
// on the first such call (GuibasStolfiDwyer_Wrapper.cpp:99-202 builds the
// reference's).  Needs triangles.
qdt_status ensure_editor(qdt_t* Q)
{
    if (Q->ed) return QDT_OK;
    qdt_status st = materialize_points(Q);
    if (st == QDT_OK) st = materialize(Q);
    if (st != QDT_OK) return st;
    std::vector<int32_t> twins;
    const int32_t* tw = nullptr;
    if (Q->ntris >= kDeviceTwinsMin && Q->eng && Q->eng->result().tris &&
        Q->eng->result().triangles == Q->ntris)
    {
        void* d = nullptr;
        if (hipMalloc(&d, 12 * Q->ntris) == hipSuccess)
        {
            uint64_t bad = 0;
            const qdt::Status s = qdt::mesh_twins(Q->eng->result().tris, (uint32_t)Q->ntris, (int*)d,
                                                  Q->eng->stream(), bad);
            if (s == qdt::Status::ok && bad == 0)
            {
                twins.resize(3 * Q->ntris);
                if (hipMemcpy(twins.data(), d, 12 * Q->ntris, hipMemcpyDeviceToHost) == hipSuccess) tw = twins.data();
            }
            (void)hipFree(d);
        }
    }
    qdt::MeshEdit* ed = new qdt::MeshEdit();
    if (!ed->build(Q->pts.p, Q->npts, Q->rep.p, Q->tris.p, Q->ntris, tw))
    {
        delete ed;
        return QDT_ERR_INTERNAL;
    }
    Q->ed = ed;
    // the mesh holds its own copies from now on
    Q->tris.reset();
    Q->rep.reset();
    Q->pts.reset();
    Q->ntris = Q->nrep = 0;
    return QDT_OK;
}

size_t cur_points(const qdt_t* Q) { return Q->ed ? Q->ed->points() : Q->npts; }
size_t cur_triangles(const qdt_t* Q)
{
    if (Q->ed) return Q->ed->triangles();
    if (Q->tris_on_device_only) return Q->eng->result().triangles;
    return Q->ntris;
}

// qdt_insert_point / qdt_insert_point_parallel (libqdt/qdt.cpp:115-140)
qdt_status insert_impl(qdt_t* Q, const qdt_vec2d& pt, size_t* index)
{
    if (!Q) return QDT_ERR_NULL_CONTEXT;
    if (!finite_pt(pt)) return QDT_ERR_NOT_FINITE;
    try
    {
        qdt_status st = materialize_points(Q);
        if (st == QDT_OK) st = materialize(Q);
        if (st != QDT_OK)
        {
            reset_context(Q);
            return st;
        }
        if (cur_points(Q) >= kMaxPoints) return QDT_ERR_TOO_MANY_POINTS;
        if (cur_triangles(Q))
        {
            // local update on the host mesh
            st = ensure_editor(Q);
            size_t idx = 0;
            if (st == QDT_OK) st = Q->ed->insert(pt, idx);
            if (st != QDT_OK)
            {
                reset_context(Q);
                return st;
            }
            if (index) *index = idx;
            return QDT_OK;
        }
        // No triangles yet (fewer than three distinct points, or all of them
        // collinear): the batch path over all points, as the reference
        // rebuilds then (GuibasStolfiDwyer_Wrapper.cpp:496-502).
        const size_t n = Q->npts + 1;
        if (!Q->pts.grow_keep(n, Q->npts)) throw std::bad_alloc();
        Q->pts.p[n - 1] = pt;
        st = ensure_engine(Q);
        if (st == QDT_OK) st = ensure_device_input(Q, n);
        if (st == QDT_OK &&
            hipMemcpyAsync(Q->d_in, Q->pts.p, n * sizeof(qdt_vec2d), hipMemcpyHostToDevice, Q->eng->stream()) !=
                hipSuccess)
            st = QDT_ERR_INTERNAL;
        if (st == QDT_OK) st = run_device(Q, n);
        if (st == QDT_OK) st = materialize(Q);
        if (st != QDT_OK)
        {
            reset_context(Q);
            return st;
        }
        Q->npts = n;
        if (index) *index = n - 1;
        return QDT_OK;
    }
    catch (const std::bad_alloc&)
    {
        reset_context(Q);
        return QDT_ERR_OUT_OF_MEMORY;
    }
    catch (...)
    {
        reset_context(Q);
        return QDT_ERR_INTERNAL;
    }
}

// Triangulation::insert_constraint (GuibasStolfiDwyer_Wrapper.cpp:1246-1262):
// checks in the order INVALID_INDEX, SAME_POINT, NO_TRIANGLES, then the
// mesh's own (CROSSES_CONSTRAINT, THROUGH_POINT).
qdt_status constraint_impl(qdt_t* Q, size_t v0, size_t v1)
{
    qdt_status st = materialize_points(Q);
    if (st == QDT_OK) st = materialize(Q);
    if (st != QDT_OK) return st;
    const size_t n = cur_points(Q);
    if (v0 >= n || v1 >= n) return QDT_ERR_INVALID_INDEX;
    const uint32_t* rep = Q->ed ? Q->ed->rep() : Q->rep.p;
    if (!Q->ed && Q->nrep < n) return QDT_ERR_INTERNAL;
    if (rep[v0] == rep[v1]) return QDT_ERR_SAME_POINT;
    if (!cur_triangles(Q)) return QDT_ERR_NO_TRIANGLES;
    st = ensure_editor(Q);
    if (st != QDT_OK) return st;
    return Q->ed->constrain(v0, v1);
}

} // namespace

extern "C" {

const char* qdt_status_string(qdt_status status)
{
    switch (status)
    {
    case QDT_OK: return "success";
    case QDT_ERR_NULL_CONTEXT: return "the context is NULL";
    case QDT_ERR_NULL_ARGUMENT: return "a required pointer argument is NULL";
    case QDT_ERR_TOO_MANY_POINTS: return "more than 100000000 points";
    case QDT_ERR_NOT_FINITE: return "a coordinate is NaN or infinite";
    case QDT_ERR_INVALID_INDEX: return "an end point index is not a point of the context";
    case QDT_ERR_SAME_POINT: return "both end points are the same point";
    case QDT_ERR_NO_TRIANGLES: return "the context has no triangles";
    case QDT_ERR_CROSSES_CONSTRAINT: return "the segment crosses an existing constraint";
    case QDT_ERR_THROUGH_POINT: return "the segment passes through another point";
    case QDT_ERR_OUT_OF_MEMORY: return "out of memory";
    case QDT_ERR_INTERNAL: return "internal error";
    }
    return "unknown status";
}

qdt_t* qdt_create(void) { return new (std::nothrow) qdt_t(); }

qdt_status qdt_execute(qdt_t* ptr, const qdt_vec2d* pts, size_t n) { return execute_host(ptr, pts, n); }

qdt_status qdt_execute_parallel(qdt_t* ptr, const qdt_vec2d* pts, size_t n) { return execute_host(ptr, pts, n); }

qdt_status qdt_insert_point(qdt_t* ptr, qdt_vec2d pt, size_t* index) { return insert_impl(ptr, pt, index); }

qdt_status qdt_insert_point_parallel(qdt_t* ptr, qdt_vec2d pt, size_t* index) { return insert_impl(ptr, pt, index); }

// libqdt/qdt.cpp:193-210: an internal failure or running out of memory
// empties the context; bad_alloc maps to QDT_ERR_OUT_OF_MEMORY, any other
// exception to QDT_ERR_INTERNAL (qdt.cpp:63-77).
qdt_status qdt_insert_constraint(qdt_t* ptr, size_t v0, size_t v1)
{
    if (!ptr) return QDT_ERR_NULL_CONTEXT;
    try
    {
        const qdt_status st = constraint_impl(ptr, v0, v1);
        if (st == QDT_ERR_INTERNAL || st == QDT_ERR_OUT_OF_MEMORY) reset_context(ptr);
        return st;
    }
    catch (const std::bad_alloc&)
    {
        reset_context(ptr);
        return QDT_ERR_OUT_OF_MEMORY;
    }
    catch (...)
    {
        reset_context(ptr);
        return QDT_ERR_INTERNAL;
    }
}

qdt_status qdt_insert_constraints_batch(qdt_t* ptr, const qdt_edge* edges, size_t num_edges, qdt_status* results)
{
    if (!ptr)
    {
        if (results)
            for (size_t i = 0; i < num_edges; ++i) results[i] = QDT_ERR_NULL_CONTEXT;
        return QDT_ERR_NULL_CONTEXT;
    }
    if (num_edges == 0) return QDT_OK;
    if (!edges)
    {
        if (results)
            for (size_t i = 0; i < num_edges; ++i) results[i] = QDT_ERR_NULL_ARGUMENT;
        return QDT_ERR_NULL_ARGUMENT;
    }
    qdt_status first = QDT_OK;
    for (size_t i = 0; i < num_edges; ++i)
    {
        const qdt_status st = qdt_insert_constraint(ptr, edges[i].v0, edges[i].v1);
        if (st == QDT_ERR_INTERNAL || st == QDT_ERR_OUT_OF_MEMORY)
        {
            if (results)
                for (size_t j = 0; j < num_edges; ++j) results[j] = st;
            return st;
        }
        if (results) results[i] = st;
        if (st != QDT_OK && first == QDT_OK) first = st;
    }
    return first;
}

size_t qdt_triangles_size(const qdt_t* ptr)
{
    if (!ptr) return 0;
    return cur_triangles(ptr);
}

size_t qdt_get_triangle(const qdt_t* ptr, size_t T[3], size_t id)
{
    if (!ptr || !T) return 0;
    qdt_t* Q = const_cast<qdt_t*>(ptr);
    if (materialize(Q) != QDT_OK) return 0;
    if (id >= cur_triangles(Q)) return 0;
    const uint32_t* t = (Q->ed ? Q->ed->tris() : Q->tris.p) + 3 * id;
    T[0] = t[0];
    T[1] = t[1];
    T[2] = t[2];
    return 1;
}

size_t qdt_points_size(const qdt_t* ptr) { return ptr ? cur_points(ptr) : 0; }

size_t qdt_get_point(const qdt_t* ptr, qdt_vec2d* pt, size_t id)
{
    if (!ptr || !pt) return 0;
    qdt_t* Q = const_cast<qdt_t*>(ptr);
    if (id >= cur_points(Q)) return 0;
    if (Q->ed)
    {
        *pt = Q->ed->pts()[id];
        return 1;
    }
    if (materialize_points(Q) != QDT_OK) return 0;
    *pt = Q->pts.p[id];
    return 1;
}

// libqdt/qdt.cpp:285-298: no triangles -> 0; else the representatives
size_t qdt_point_is_used(const qdt_t* ptr, size_t id)
{
    if (!ptr) return 0;
    qdt_t* Q = const_cast<qdt_t*>(ptr);
    if (Q->ed) return id < Q->ed->points() && Q->ed->triangles() && Q->ed->rep()[id] == id ? 1 : 0;
    if (materialize(Q) != QDT_OK) return 0;
    if (id >= Q->npts || id >= Q->nrep) return 0;
    if (Q->ntris == 0) return 0;
    return Q->rep.p[id] == id ? 1 : 0;
}

void qdt_destroy(qdt_t* ptr) { delete ptr; }

// ---- GPU extensions ------------------------------------------------------

qdt_status qdt_gpu_execute_device(qdt_t* ptr, const qdt_vec2d* d_pts, size_t n)
{
    if (!ptr) return QDT_ERR_NULL_CONTEXT;
    qdt_t* Q = ptr;
    try
    {
        qdt_status st = batch_checks(d_pts, n);
        reset_context(Q);
        if (st != QDT_OK || n == 0) return st;
        st = ensure_engine(Q);
        if (st == QDT_OK) st = ensure_device_input(Q, n);
        if (st == QDT_OK && hipMemcpyAsync(Q->d_in, d_pts, n * sizeof(qdt_vec2d), hipMemcpyDeviceToDevice,
                                           Q->eng->stream()) != hipSuccess)
            st = QDT_ERR_INTERNAL;
        if (st == QDT_OK) st = run_device(Q, n);
        if (st != QDT_OK)
        {
            reset_context(Q);
            return st;
        }
        Q->npts = n;
        Q->pts_on_device_only = true;
        return QDT_OK;
    }
    catch (...)
    {
        reset_context(Q);
        return QDT_ERR_OUT_OF_MEMORY;
    }
}

size_t qdt_gpu_device_triangles(const qdt_t* ptr, const uint32_t** d_tris)
{
    if (!ptr || !ptr->eng || ptr->ed || (!ptr->tris_on_device_only && ptr->ntris == 0))
    {
        if (d_tris) *d_tris = nullptr;
        return 0;
    }
    const qdt::DeviceResult& r = ptr->eng->result();
    if (d_tris) *d_tris = r.tris;
    return r.triangles;
}

size_t qdt_gpu_copy_triangles(const qdt_t* ptr, uint32_t* out)
{
    if (!ptr || !out) return 0;
    qdt_t* Q = const_cast<qdt_t*>(ptr);
    if (Q->ed)
    {
        std::memcpy(out, Q->ed->tris(), 12 * Q->ed->triangles());
        return Q->ed->triangles();
    }
    if (Q->tris_on_device_only)
    {
        const qdt::DeviceResult& r = Q->eng->result();
        if (r.triangles && hipMemcpy(out, r.tris, 12 * (size_t)r.triangles, hipMemcpyDeviceToHost) != hipSuccess)
            return 0;
        return r.triangles;
    }
    std::memcpy(out, Q->tris.p, 12 * Q->ntris);
    return Q->ntris;
}

int qdt_gpu_get_stats(const qdt_t* ptr, qdt_gpu_stats* out)
{
    if (!ptr || !out) return 0;
    *out = ptr->stats;
    return 1;
}

void qdt_gpu_set_profiling(qdt_t* ptr, int enable)
{
    if (!ptr) return;
    ptr->profiling = enable != 0;
    if (ptr->eng) ptr->eng->set_profiling(ptr->profiling);
}

void qdt_gpu_set_seams(qdt_t* ptr, int enable)
{
    if (!ptr) return;
    ptr->seams = enable != 0;
    if (ptr->eng) ptr->eng->set_seams(ptr->seams);
}

void qdt_gpu_set_cells(qdt_t* ptr, int enable)
{
    if (!ptr) return;
    ptr->cells = enable != 0;
    if (ptr->eng) ptr->eng->set_cells(ptr->cells);
}

void qdt_gpu_set_test_limits(qdt_t* ptr, size_t device_bytes, size_t spare_records)
{
    if (!ptr) return;
    ptr->mem_cap = device_bytes;
    ptr->spare_limit = spare_records;
    if (ptr->eng) ptr->eng->set_limits(device_bytes, spare_records);
}

int qdt_gpu_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

namespace
{

void copy_report(const qdt::VerifyReport& r, qdt_gpu_verify_report* o)
{
    o->triangles = r.triangles;
    o->boundary_edges = r.boundary_edges;
    o->vertices = r.vertices;
    o->bad_index = r.bad_index;
    o->not_representative = r.not_representative;
    o->bad_orientation = r.bad_orientation;
    o->bad_edges = r.bad_edges;
    o->non_delaunay = r.non_delaunay;
    o->concave_boundary = r.concave_boundary;
    o->unused_points = r.unused_points;
    o->off_line = r.off_line;
    o->euler_ok = r.euler_ok ? 1 : 0;
    o->ok = r.ok ? 1 : 0;
    o->device_ms = r.device_ms;
}

qdt_status status_of(qdt::Status s)
{
    return s == qdt::Status::ok ? QDT_OK
                                : (s == qdt::Status::out_of_memory ? QDT_ERR_OUT_OF_MEMORY : QDT_ERR_INTERNAL);
}

} // namespace

qdt_status qdt_gpu_verify_triangles(const qdt_vec2d* d_pts, size_t n, const uint32_t* d_tris, size_t T,
                                    const uint32_t* d_rep, qdt_gpu_verify_report* out)
{
    if (!out || (n && !d_pts) || (T && !d_tris)) return QDT_ERR_NULL_ARGUMENT;
    if (n > 100000000u || T > 400000000u) return QDT_ERR_TOO_MANY_POINTS;
    try
    {
        qdt::VerifyReport r;
        const qdt::Status s = qdt::verify_triangulation((const double*)d_pts, (uint32_t)n, d_tris, (uint32_t)T,
                                                        d_rep, nullptr, r);
        if (s == qdt::Status::ok) copy_report(r, out);
        return status_of(s);
    }
    catch (...)
    {
        return QDT_ERR_OUT_OF_MEMORY;
    }
}

qdt_status qdt_gpu_build_mesh(const uint32_t* d_tris, size_t T, size_t n, int32_t* d_he, int32_t* d_vert,
                              int32_t* d_orig_to_local, qdt_gpu_mesh_info* out)
{
    if (!out || (T && (!d_tris || !d_he || !d_vert)) || (n && !d_orig_to_local)) return QDT_ERR_NULL_ARGUMENT;
    if (n > 100000000u || T > 400000000u) return QDT_ERR_TOO_MANY_POINTS;
    try
    {
        qdt::MeshReport r;
        const qdt::Status s = qdt::build_mesh(d_tris, (uint32_t)T, (uint32_t)n, (int4*)d_he, (int2*)d_vert,
                                              (int*)d_orig_to_local, nullptr, r);
        if (s != qdt::Status::ok) return status_of(s);
        out->vertices = r.vertices;
        out->halfedges = r.halfedges;
        out->boundary = r.boundary;
        out->unmatched = r.unmatched;
        out->device_ms = r.device_ms;
        return r.bad_index ? QDT_ERR_INVALID_INDEX : QDT_OK;
    }
    catch (...)
    {
        return QDT_ERR_OUT_OF_MEMORY;
    }
}

qdt_status qdt_gpu_verify(qdt_t* ptr, qdt_gpu_verify_report* out)
{
    if (!ptr) return QDT_ERR_NULL_CONTEXT;
    if (!out) return QDT_ERR_NULL_ARGUMENT;
    qdt_t* Q = ptr;
    try
    {
        qdt::VerifyReport r;
        if (cur_points(Q) == 0 || !Q->eng)
        {
            r.ok = r.euler_ok = true;
            copy_report(r, out);
            return QDT_OK;
        }
        qdt::Status s;
        if (Q->tris_on_device_only)
        {
            // the last run's results are still only on the device
            const qdt::DeviceResult& d = Q->eng->result();
            s = qdt::verify_triangulation(Q->d_in, (uint32_t)Q->npts, d.tris, d.triangles, d.rep,
                                          Q->eng->stream(), r);
        }
        else
        {
            // the host store (the device workspace may have been released),
            // or the host mesh after insertions and constraints
            qdt_status hs = materialize_points(Q);
            if (hs != QDT_OK) return hs;
            const qdt_vec2d* hp = Q->ed ? Q->ed->pts() : Q->pts.p;
            const uint32_t* ht = Q->ed ? Q->ed->tris() : Q->tris.p;
            const uint32_t* hr = Q->ed ? Q->ed->rep() : Q->rep.p;
            void *dp = nullptr, *dt = nullptr, *dr = nullptr;
            struct Free
            {
                void** p[3];
                ~Free()
                {
                    for (void** q : p)
                        if (*q) (void)hipFree(*q);
                }
            } fr{ { &dp, &dt, &dr } };
            const size_t np = cur_points(Q), nt = cur_triangles(Q);
            if (hipMalloc(&dp, 16 * np) != hipSuccess || hipMalloc(&dt, 12 * std::max<size_t>(nt, 1)) != hipSuccess ||
                hipMalloc(&dr, 4 * np) != hipSuccess)
                return QDT_ERR_OUT_OF_MEMORY;
            if (hipMemcpy(dp, hp, 16 * np, hipMemcpyHostToDevice) != hipSuccess ||
                (nt && hipMemcpy(dt, ht, 12 * nt, hipMemcpyHostToDevice) != hipSuccess) ||
                hipMemcpy(dr, hr, 4 * np, hipMemcpyHostToDevice) != hipSuccess)
                return QDT_ERR_INTERNAL;
            s = qdt::verify_triangulation((const double*)dp, (uint32_t)np, (const uint32_t*)dt, (uint32_t)nt,
                                          (const uint32_t*)dr, Q->eng->stream(), r);
        }
        if (s == qdt::Status::ok) copy_report(r, out);
        return status_of(s);
    }
    catch (...)
    {
        return QDT_ERR_OUT_OF_MEMORY;
    }
}

void qdt_gpu_release_workspace(qdt_t* ptr)
{
    if (!ptr || !ptr->eng) return;
    qdt_t* Q = ptr;
    if (materialize(Q) != QDT_OK) reset_context(Q);
    if (materialize_points(Q) != QDT_OK) reset_context(Q);
    Q->eng->release();
}

// ---- batched form ----------------------------------------------------------

} // extern "C"

struct qdt_gpu_batch
{
    qdt::Engine* eng = nullptr;
    double* d_in = nullptr;
    size_t d_in_cap = 0;
    size_t k = 0, n = 0;
    Pinned<uint32_t> tris; // host copy of the grouped triangles (lazy)
    bool tris_host = false;
    qdt_gpu_stats stats{};
    bool profiling = false;
    bool seams = true;
    bool graphs = true;
    int cell_points = 0, merge_lanes = 0; // 0: autotuned
    qdt::TuneChoice tune;
    ~qdt_gpu_batch()
    {
        delete eng;
        if (d_in) (void)hipFree(d_in);
    }
};

namespace
{

void reset_batch(qdt_gpu_batch* B)
{
    B->k = B->n = 0;
    B->tris_host = false;
    B->stats = qdt_gpu_stats{};
}

qdt_status batch_engine(qdt_gpu_batch* B)
{
    if (B->eng) return QDT_OK;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return QDT_ERR_INTERNAL;
    B->eng = new (std::nothrow) qdt::Engine();
    if (!B->eng) return QDT_ERR_OUT_OF_MEMORY;
    B->eng->set_profiling(B->profiling);
    B->eng->set_seams(B->seams);
    B->eng->set_graphs(B->graphs);
    if (!B->eng->stream())
    {
        delete B->eng;
        B->eng = nullptr;
        return QDT_ERR_INTERNAL;
    }
    return QDT_OK;
}

qdt_status batch_offsets(const size_t* offsets, size_t k)
{
    if (k == 0) return QDT_OK;
    if (!offsets) return QDT_ERR_NULL_ARGUMENT;
    if (offsets[0] != 0) return QDT_ERR_INVALID_INDEX;
    for (size_t i = 0; i < k; ++i)
        if (offsets[i + 1] < offsets[i]) return QDT_ERR_INVALID_INDEX;
    if (offsets[k] > kMaxPoints) return QDT_ERR_TOO_MANY_POINTS;
    if (k > (size_t)INT32_MAX) return QDT_ERR_TOO_MANY_POINTS;
    return QDT_OK;
}

// pts: host (device == false) or device memory
qdt_status batch_run(qdt_gpu_batch* B, const qdt_vec2d* pts, const size_t* offsets, size_t k, bool device)
{
    if (!B) return QDT_ERR_NULL_CONTEXT;
    try
    {
        reset_batch(B);
        qdt_status st = batch_offsets(offsets, k);
        if (st != QDT_OK || k == 0) return st;
        const size_t n = offsets[k];
        if (n && !pts) return QDT_ERR_NULL_ARGUMENT;
        st = batch_engine(B);
        if (st != QDT_OK) return st;
        const double* src = (const double*)pts;
        if (!device && n)
        {
            const size_t bytes = n * sizeof(qdt_vec2d);
            if (bytes > B->d_in_cap)
            {
                if (B->d_in) (void)hipFree(B->d_in);
                B->d_in = nullptr;
                B->d_in_cap = 0;
                void* p = nullptr;
                const hipError_t e = hipMalloc(&p, bytes);
                if (e != hipSuccess) return e == hipErrorOutOfMemory ? QDT_ERR_OUT_OF_MEMORY : QDT_ERR_INTERNAL;
                B->d_in = (double*)p;
                B->d_in_cap = bytes;
            }
            if (hipMemcpy(B->d_in, pts, bytes, hipMemcpyHostToDevice) != hipSuccess) return QDT_ERR_INTERNAL;
            src = B->d_in;
        }
        qdt::Status s = qdt::Status::ok;
        if (B->cell_points && B->merge_lanes)
        {
            B->eng->set_tuning(B->cell_points, B->merge_lanes);
            B->tune = qdt::TuneChoice();
            B->tune.cell_points = B->eng->cell_points();
            B->tune.merge_lanes = B->eng->merge_lanes();
        }
        else
            s = qdt::tune_batch(*B->eng, src, offsets, (int)k, B->tune);
        if (s == qdt::Status::ok) s = B->eng->run_batch(src, offsets, (int)k);
        if (s != qdt::Status::ok)
        {
            reset_batch(B);
            copy_stats(B->eng->stats(), B->stats); // the failed run's counters and error flags
            return map_status(s);
        }
        copy_stats(B->eng->stats(), B->stats);
        B->stats.tune_source = (uint64_t)B->tune.source;
        B->stats.tune_ms = B->tune.device_ms;
        B->k = k;
        B->n = n;
        return QDT_OK;
    }
    catch (...)
    {
        reset_batch(B);
        return QDT_ERR_OUT_OF_MEMORY;
    }
}

bool batch_materialize(const qdt_gpu_batch* Bc)
{
    qdt_gpu_batch* B = const_cast<qdt_gpu_batch*>(Bc);
    if (B->tris_host) return true;
    const qdt::BatchResult& r = B->eng->batch();
    if (!B->tris.reserve(3 * (size_t)r.triangles + 3)) return false;
    if (r.triangles && hipMemcpy(B->tris.p, r.tris, 12 * (size_t)r.triangles, hipMemcpyDeviceToHost) != hipSuccess)
        return false;
    B->tris_host = true;
    return true;
}

const qdt::BatchResult* batch_result(const qdt_gpu_batch* B, size_t set)
{
    if (!B || !B->eng || set >= B->k) return nullptr;
    const qdt::BatchResult& r = B->eng->batch();
    if ((size_t)r.k != B->k) return nullptr;
    return &r;
}

} // namespace

extern "C" {

qdt_gpu_batch* qdt_gpu_batch_create(void) { return new (std::nothrow) qdt_gpu_batch(); }

void qdt_gpu_batch_destroy(qdt_gpu_batch* b) { delete b; }

qdt_status qdt_gpu_batch_execute(qdt_gpu_batch* b, const qdt_vec2d* pts, const size_t* offsets, size_t k)
{
    return batch_run(b, pts, offsets, k, false);
}

qdt_status qdt_gpu_batch_execute_device(qdt_gpu_batch* b, const qdt_vec2d* d_pts, const size_t* offsets, size_t k)
{
    return batch_run(b, d_pts, offsets, k, true);
}

size_t qdt_gpu_batch_sets(const qdt_gpu_batch* b) { return b ? b->k : 0; }

qdt_status qdt_gpu_batch_set_status(const qdt_gpu_batch* b, size_t set)
{
    if (!b) return QDT_ERR_NULL_CONTEXT;
    const qdt::BatchResult* r = batch_result(b, set);
    if (!r) return QDT_ERR_INVALID_INDEX;
    return (qdt_status)r->status[set];
}

size_t qdt_gpu_batch_triangles_size(const qdt_gpu_batch* b, size_t set)
{
    const qdt::BatchResult* r = batch_result(b, set);
    return r ? (size_t)r->tri_count[set] : 0;
}

size_t qdt_gpu_batch_unique_points(const qdt_gpu_batch* b, size_t set)
{
    const qdt::BatchResult* r = batch_result(b, set);
    return r ? (size_t)r->unique[set] : 0;
}

size_t qdt_gpu_batch_get_triangle(const qdt_gpu_batch* b, size_t set, size_t T[3], size_t id)
{
    const qdt::BatchResult* r = batch_result(b, set);
    if (!r || !T || id >= r->tri_count[set]) return 0;
    if (!batch_materialize(b)) return 0;
    const uint32_t* t = b->tris.p + 3 * (r->tri_start[set] + id);
    T[0] = t[0];
    T[1] = t[1];
    T[2] = t[2];
    return 1;
}

int qdt_gpu_batch_set_range(const qdt_gpu_batch* b, size_t set, size_t* first, size_t* count)
{
    const qdt::BatchResult* r = batch_result(b, set);
    if (!r) return 0;
    if (first) *first = (size_t)r->tri_start[set];
    if (count) *count = (size_t)r->tri_count[set];
    return 1;
}

size_t qdt_gpu_batch_copy_triangles(const qdt_gpu_batch* b, uint32_t* out)
{
    if (!b || !b->eng || !b->k || !out) return 0;
    const qdt::BatchResult& r = b->eng->batch();
    if (b->tris_host)
    {
        std::memcpy(out, b->tris.p, 12 * (size_t)r.triangles);
        return (size_t)r.triangles;
    }
    if (r.triangles && hipMemcpy(out, r.tris, 12 * (size_t)r.triangles, hipMemcpyDeviceToHost) != hipSuccess)
        return 0;
    return (size_t)r.triangles;
}

size_t qdt_gpu_batch_device_triangles(const qdt_gpu_batch* b, const uint32_t** d_tris)
{
    if (!b || !b->eng || !b->k)
    {
        if (d_tris) *d_tris = nullptr;
        return 0;
    }
    const qdt::BatchResult& r = b->eng->batch();
    if (d_tris) *d_tris = r.tris;
    return (size_t)r.triangles;
}

size_t qdt_gpu_batch_copy_rep(const qdt_gpu_batch* b, uint32_t* out)
{
    if (!b || !b->eng || !b->k || !out) return 0;
    const qdt::BatchResult& r = b->eng->batch();
    if (r.n && hipMemcpy(out, r.rep, 4 * r.n, hipMemcpyDeviceToHost) != hipSuccess) return 0;
    return r.n;
}

int qdt_gpu_batch_get_stats(const qdt_gpu_batch* b, qdt_gpu_stats* out)
{
    if (!b || !out) return 0;
    *out = b->stats;
    return 1;
}

void qdt_gpu_batch_set_profiling(qdt_gpu_batch* b, int enable)
{
    if (!b) return;
    b->profiling = enable != 0;
    if (b->eng) b->eng->set_profiling(b->profiling);
}

void qdt_gpu_batch_set_seams(qdt_gpu_batch* b, int enable)
{
    if (!b) return;
    b->seams = enable != 0;
    if (b->eng) b->eng->set_seams(b->seams);
}

void qdt_gpu_batch_set_graphs(qdt_gpu_batch* b, int enable)
{
    if (!b) return;
    b->graphs = enable != 0;
    if (b->eng) b->eng->set_graphs(b->graphs);
}

void qdt_gpu_batch_set_tuning(qdt_gpu_batch* b, int cell_points, int merge_lanes)
{
    if (!b) return;
    b->cell_points = cell_points > 0 ? cell_points : 0;
    b->merge_lanes = merge_lanes > 0 ? merge_lanes : 0;
}

} // extern "C"
