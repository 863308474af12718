// ref_shim.cpp -- exports the reference's own predicates (libqdt/robust.hxx)
// so that the oracle's restatement can be checked against them.  Compiled only
// into oracle/_ref/libqdt_ref.so, together with the reference's libqdt/*.cpp
// taken from where they lie; test infrastructure, never part of the product.
#include "robust.hxx"

extern "C" double ref_orient2d(double ax, double ay, double bx, double by, double cx, double cy)
{
    const wrapper::Point a = { ax, ay }, b = { bx, by }, c = { cx, cy };
    return wrapper::robust::orient2d(a, b, c);
}

extern "C" double ref_incircle(double ax, double ay, double bx, double by, double cx, double cy,
                               double dx, double dy)
{
    const wrapper::Point a = { ax, ay }, b = { bx, by }, c = { cx, cy }, d = { dx, dy };
    return wrapper::robust::incircle(a, b, c, d);
}

// Bulk read of a reference context's triangles (test speed only): reads the
// public result fields of the context type declared in libqdt/wrapper.hxx.
#include "qdt.h"
#include "wrapper.hxx"
#include <cstdint>

extern "C" size_t ref_copy_triangles(const qdt_t* q, uint32_t* out)
{
    const wrapper::Triangulation* T = reinterpret_cast<const wrapper::Triangulation*>(q);
    const size_t n = T->triangles.size();
    if (out)
        for (size_t i = 0; i < n; ++i)
            for (int k = 0; k < 3; ++k) out[3 * i + k] = static_cast<uint32_t>(T->triangles[i][k]);
    return n;
}

// Many independent sets on a pool of host threads, each thread with its own
// reference context (qdt_execute, serial path, per set): the CPU baseline of
// the batched form.  Returns the wall seconds of the whole pool, or -1.
#include <atomic>
#include <chrono>
#include <thread>
#include <vector>

extern "C" double ref_batch_seconds(const double* xy, const uint64_t* offs, size_t k, int threads,
                                    uint64_t* total_tris)
{
    std::atomic<size_t> next{ 0 };
    std::atomic<uint64_t> tris{ 0 };
    std::atomic<int> bad{ 0 };
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&] {
            qdt_t* q = qdt_create();
            for (;;)
            {
                const size_t i = next.fetch_add(1);
                if (i >= k) break;
                const qdt_vec2d* p = reinterpret_cast<const qdt_vec2d*>(xy + 2 * offs[i]);
                if (qdt_execute(q, p, (size_t)(offs[i + 1] - offs[i])) != QDT_OK) bad.fetch_add(1);
                tris.fetch_add(qdt_triangles_size(q));
            }
            qdt_destroy(q);
        });
    for (std::thread& th : pool) th.join();
    const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (total_tris) *total_tris = tris.load();
    return bad.load() ? -1.0 : dt;
}

// Seconds of the reference's half-edge mesh build (Triangulation::ensure_mesh_
// or ensure_mesh_parallel_, GuibasStolfiDwyer_Wrapper.cpp:99-398; public
// members of the context type) on a context that already holds the batch
// result of xy (qdt_execute_parallel, not timed): the CPU baseline of the
// device mesh build.  halfedges receives 3 T; -1 on failure.
extern "C" double ref_mesh_seconds(const double* xy, size_t n, int parallel, uint64_t* halfedges)
{
    qdt_t* q = qdt_create();
    if (!q) return -1.0;
    if (qdt_execute_parallel(q, reinterpret_cast<const qdt_vec2d*>(xy), n) != QDT_OK)
    {
        qdt_destroy(q);
        return -1.0;
    }
    wrapper::Triangulation* T = reinterpret_cast<wrapper::Triangulation*>(q);
    const auto t0 = std::chrono::steady_clock::now();
    if (parallel)
        T->ensure_mesh_parallel_();
    else
        T->ensure_mesh_();
    const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (halfedges) *halfedges = 3 * (uint64_t)T->triangles.size();
    qdt_destroy(q);
    return dt;
}

// Per-set results of the reference for many independent sets (the batched
// and served workloads' parity fixtures): set i = points [offs[i],
// offs[i+1]) through qdt_execute on its own context, on `threads` host
// threads.  counts[i] = triangles, hashes[i] = FNV-1a-64 of the canonical
// triangle list (each triangle rotated so that its smallest index comes
// first, the list sorted, every index written as 8 little-endian bytes).
// Returns 0, or -1 when a set fails.
#include <algorithm>
#include <array>

static uint64_t canon_fnv(const qdt_t* q)
{
    const wrapper::Triangulation* T = reinterpret_cast<const wrapper::Triangulation*>(q);
    std::vector<std::array<uint64_t, 3>> c;
    c.reserve(T->triangles.size());
    for (const auto& t : T->triangles)
    {
        int r = 0;
        if (t[1] < t[r]) r = 1;
        if (t[2] < t[r]) r = 2;
        c.push_back({ (uint64_t)t[r], (uint64_t)t[(r + 1) % 3], (uint64_t)t[(r + 2) % 3] });
    }
    std::sort(c.begin(), c.end());
    uint64_t h = 0xcbf29ce484222325ull;
    for (const auto& t : c)
        for (uint64_t v : t)
            for (int b = 0; b < 8; ++b)
            {
                h ^= (v >> (8 * b)) & 0xff;
                h *= 0x100000001b3ull;
            }
    return h;
}

extern "C" int ref_batch_hashes(const double* xy, const uint64_t* offs, size_t k, int threads, uint64_t* counts,
                                uint64_t* hashes)
{
    std::atomic<size_t> next{ 0 };
    std::atomic<int> bad{ 0 };
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&] {
            qdt_t* q = qdt_create();
            for (;;)
            {
                const size_t i = next.fetch_add(1);
                if (i >= k) break;
                const qdt_vec2d* p = reinterpret_cast<const qdt_vec2d*>(xy + 2 * offs[i]);
                if (qdt_execute(q, p, (size_t)(offs[i + 1] - offs[i])) != QDT_OK) bad.fetch_add(1);
                counts[i] = qdt_triangles_size(q);
                hashes[i] = canon_fnv(q);
            }
            qdt_destroy(q);
        });
    for (std::thread& th : pool) th.join();
    return bad.load() ? -1 : 0;
}
