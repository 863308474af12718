// engine.hip -- the GPU batch Delaunay pipeline for gfx950.
//
// Stages (one HIP stream, no host work between them except reading m):
//   1. keys:      order-preserving 64-bit keys of x and y, finiteness check
//   2. sort:      stable LSD radix sorts by y then x -> (x, y, index) order
//   3. dedupe:    runs of equal coordinates -> distinct points P0 (frame-0
//                 order), lowest index ID0, rep[] for every input index
//                 (libqdt/builder.hxx:356-390)
//   4. frame 1:   stable sort of the reversed frame-0 order by y -> the
//                 (y asc, x desc) order of builder.hxx:75
//   5. kd levels: Dwyer's alternating median cuts (builder.hxx:303-334),
//                 level by level: a positional split of the array sorted in
//                 the cut's frame and a stable two-way partition of the
//                 array sorted in the other frame, so both stay sorted
//   6. leaves:    2- and 3-point leaves (builder.hxx:177-201)
//   7. merges:    Guibas-Stolfi merges bottom-up, one thread per kd node,
//                 one launch per level (builder.hxx:205-301)
//   8. extract:   faces of the quad-edge store with a filtered orient2d;
//                 undecided tests go through a queue to an exact kernel;
//                 compaction in record order (builder.hxx:98-116)
// Topology equals the reference's for every input: the same cuts, leaves,
// merges and hull walks run with exact predicates.
#include "engine.hpp"
#include "primitives.hpp"
#include "quadedge.hpp"
#include "seam.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

namespace qdt
{

namespace
{

constexpr int kMergeBlock = 64;

#define QDT_HIP(x)                                                                                       \
    do                                                                                                   \
    {                                                                                                    \
        hipError_t e_ = (x);                                                                             \
        if (e_ != hipSuccess) return e_ == hipErrorOutOfMemory ? Status::out_of_memory : Status::internal; \
    } while (0)

#define QDT_TRY(x)                                                                                       \
    do                                                                                                   \
    {                                                                                                    \
        const Status st_try_ = (x);                                                                      \
        if (st_try_ != Status::ok) return st_try_;                                                       \
    } while (0)

inline uint32_t cdiv(size_t a, size_t b) { return (uint32_t)((a + b - 1) / b); }

__device__ __forceinline__ uint64_t order_key(double v)
{
    // -0.0 compares equal to +0.0, so both map to the key of +0.0
    const uint64_t b = v == 0.0 ? 0ull : (uint64_t)__double_as_longlong(v);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

// counters[]: 0 orient, 1 orient exact, 2 incircle, 3 incircle exact, 4 extract queue
__device__ __forceinline__ void flush_counts(unsigned long long* counters, const PredCount& pc)
{
    unsigned long long v[4] = { pc.orient, pc.orient_exact, pc.incircle, pc.incircle_exact };
#pragma unroll
    for (int i = 0; i < 4; ++i)
    {
        unsigned long long x = v[i];
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
        if (__lane_id() == 0 && x) atomicAdd(&counters[i], x);
    }
}

// Sum of v over the wavefront added to *c by one atomic (every lane of the
// wavefront must call it; lanes with nothing to add pass 0).
__device__ __forceinline__ void wave_add(unsigned long long* c, unsigned long long v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    if (__lane_id() == 0 && v) atomicAdd(c, v);
}

// Wavefront-aggregated fetch-and-add: one atomic per wavefront, each lane
// gets the old value plus the sum of the lower lanes' v (the whole wavefront
// must call it; lanes with nothing to take pass 0).
template <class T>
__device__ __forceinline__ T wave_fetch_add(T* c, T v)
{
    const int lane = (int)__lane_id();
    T x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1)
    {
        const T y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    const T total = __shfl(x, 63, 64);
    T base = 0;
    if (lane == 63 && total) base = atomicAdd(c, total);
    base = __shfl(base, 63, 64);
    return base + x - v;
}

// edges created by the merge kernel (for its algorithmic bytes)
__device__ __forceinline__ void flush_connects(unsigned long long* counters, const PredCount& pc)
{
    unsigned long long x = pc.connects;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
    if (__lane_id() == 0 && x) atomicAdd(&counters[13], x);
}

// ---------------------------------------------------------------------------
// 1-4: keys, sorts, dedupe, frame-1 order
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(kBlock) void k_keys(const double2* pts, uint32_t n, uint64_t* kx, uint64_t* ky,
                                                 uint32_t* idx, int* err)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const double2 p = pts[i];
    if (!isfinite(p.x) || !isfinite(p.y)) atomicOr(err, 1);
    kx[i] = order_key(p.x);
    ky[i] = order_key(p.y);
    idx[i] = i;
}

__global__ __launch_bounds__(kBlock) void k_gather_u64(const uint32_t* idx, const uint64_t* src, uint32_t n,
                                                       uint64_t* dst)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) dst[i] = src[idx[i]];
}

// Sorted position p starts a new distinct point when its order keys (x, y)
// or, with several sets, its set rank differ from position p-1's (F[p]).
// The point is gathered here once, in sorted order, for the write that
// follows (Pg).
__global__ __launch_bounds__(kBlock) void k_dedupe_flags(const double2* pts, const uint32_t* idx,
                                                         const uint64_t* rank, uint32_t n, uint8_t* F, double2* Pg)
{
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t q = p < n ? p : n - 1; // every lane takes part in the shuffles
    const double2 a = pts[idx[q]];
    // position p-1's point: the lane below's, or loaded by lane 0
    const int lane = (int)(threadIdx.x & 63);
    double2 b;
    b.x = __shfl_up(a.x, 1);
    b.y = __shfl_up(a.y, 1);
    if (lane == 0 && q > 0) b = pts[idx[q - 1]];
    if (p >= n) return;
    Pg[p] = a;
    uint32_t f = 1;
    if (p > 0)
    {
        f = (order_key(a.x) != order_key(b.x) || order_key(a.y) != order_key(b.y) ||
             (rank && rank[p] != rank[p - 1]))
                ? 1u
                : 0u;
    }
    F[p] = (uint8_t)f;
}

struct LoadU32
{
    const uint32_t* in;
    __device__ uint32_t operator()(uint32_t i) const { return in[i]; }
};

struct StoreU32
{
    uint32_t* out;
    __device__ void operator()(uint32_t i, uint32_t v) const { out[i] = v; }
};

__global__ __launch_bounds__(kBlock) void k_dedupe_write(const double2* Pg, const uint32_t* idx, const uint8_t* F,
                                                         const uint64_t* rank, const uint32_t* dist, uint32_t n,
                                                         double2* P0, uint32_t* ID0, uint32_t* vrank, uint32_t* didx)
{
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= n) return;
    if (F[p])
    {
        const uint32_t d = dist[p];
        const uint32_t i = idx[p];
        P0[d] = Pg[p];
        ID0[d] = i;
        if (rank) vrank[d] = (uint32_t)rank[p];
        if (didx) didx[i] = d;
    }
}

// After the sort by y (ties by input index): keeps that order for frame 1.
__global__ __launch_bounds__(kBlock) void k_ytie(const uint32_t* idx, uint32_t n, uint32_t* perm)
{
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p < n) perm[p] = idx[p];
}

// Frame 1 orders points of equal y by x descending, which the y sort's
// index order does not give: the first thread of every run of equal y keys
// re-sorts its run in perm by x descending (insertion sort, stable, so
// copies keep their index order).  A run longer than kYRunMax is left alone
// and flagged (tie): frame 1 then comes from a full sort.
constexpr uint32_t kYRunMax = 64;
__global__ __launch_bounds__(kBlock) void k_yfix(const uint64_t* ysorted, const uint64_t* kx, uint32_t n,
                                                 uint32_t* perm, unsigned long long* tie)
{
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p + 1 >= n) return;
    const uint64_t y = ysorted[p];
    if (ysorted[p + 1] != y || (p > 0 && ysorted[p - 1] == y)) return; // not the first of a run
    uint32_t q = p + 2;
    while (q < n && q - p <= kYRunMax && ysorted[q] == y) ++q;
    if (q - p > kYRunMax)
    {
        *tie = 1;
        return;
    }
    for (uint32_t a = p + 1; a < q; ++a)
    {
        const uint32_t v = perm[a];
        const uint64_t xv = kx[v];
        uint32_t b = a;
        while (b > p && kx[perm[b - 1]] < xv)
        {
            perm[b] = perm[b - 1];
            --b;
        }
        perm[b] = v;
    }
}

// frame-1 order without a sort: the y order's representatives (a point's
// lowest-index copy) in that order, as distinct ids
// distinct-point id of every input index that represents its point (the
// first of its duplicates), ~0 for the others; in input order, so that the
// scan below gathers one word per position
__global__ __launch_bounds__(kBlock) void k_rep_id(const uint32_t* rep, const uint32_t* didx, uint32_t n, uint32_t* w)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) w[i] = rep[i] == i ? didx[i] : ~0u;
}

// w in y order, gathered once by position (g) for both passes of the scan
__global__ __launch_bounds__(kBlock) void k_rep_gather(const uint32_t* perm, const uint32_t* w, uint32_t n,
                                                       uint32_t* g)
{
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p < n) g[p] = w[perm[p]];
}

struct RepLoad
{
    const uint32_t* g;
    __device__ uint32_t operator()(uint32_t p) const { return g[p] != ~0u ? 1u : 0u; }
};

struct RepWrite
{
    const uint32_t* g;
    uint32_t* out;
    __device__ uint32_t load(uint32_t p) const { return g[p]; }
    __device__ void store(uint32_t, uint32_t pre, uint32_t v) const
    {
        if (v != ~0u) out[pre] = v;
    }
};

// ---- several point sets in one run ----

// set of input index i: the last s with soff[s] <= i
__device__ __forceinline__ int set_of(const uint32_t* soff, int k, uint32_t i)
{
    int lo = 0, hi = k - 1;
    while (lo < hi)
    {
        const int mid = (lo + hi + 1) >> 1;
        if (soff[mid] <= i)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(kBlock) void k_check_sets(const double2* pts, uint32_t n, const uint32_t* soff, int k,
                                                       uint32_t* setbad)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const double2 p = pts[i];
    if (!isfinite(p.x) || !isfinite(p.y)) atomicOr(&setbad[set_of(soff, k, i)], 1u);
}

// rank of every input point's set (the sets' order in the kd forest)
__global__ __launch_bounds__(kBlock) void k_input_rank(uint32_t n, const uint32_t* soff, int k, const uint32_t* srank,
                                                       uint32_t* rin)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) rin[i] = srank[set_of(soff, k, i)];
}

__global__ __launch_bounds__(kBlock) void k_gather_rank(const uint32_t* idx, const uint32_t* rank, uint32_t n,
                                                        uint64_t* key)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) key[i] = rank[idx[i]];
}

// (set rank, input index) of every position of an order of input indices
__global__ __launch_bounds__(kBlock) void k_rank_pairs(const uint32_t* idx, const uint32_t* rank, uint32_t n,
                                                       uint64_t* key, uint32_t* val)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t v = idx[i];
    key[i] = rank[v];
    val[i] = v;
}

// first distinct point of every rank; rpos[r] = first sorted position of rank r
__global__ __launch_bounds__(kBlock) void k_vstart(const uint32_t* dist, const uint32_t* rpos, int nr, uint32_t n,
                                                   const uint32_t* m_total, uint32_t* vstart)
{
    const int r = blockIdx.x * kBlock + threadIdx.x;
    if (r > nr) return;
    const uint32_t p = rpos[r];
    vstart[r] = p < n ? dist[p] : *m_total;
}

// rank of every triangle (by its first vertex's input index); T on the device
__global__ __launch_bounds__(kBlock) void k_tri_rank(const uint32_t* tris, const uint32_t* dT, const uint32_t* rin,
                                                     uint32_t* trank)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < *dT) trank[i] = rin[tris[3 * i]];
}

// lower bounds of ranks 0..nr in a rank-sorted key array of min(*dlen, len)
// keys (dlen may be null)
template <class K>
__global__ __launch_bounds__(kBlock) void k_rank_bounds(const K* keys, const uint32_t* dlen, uint32_t len, int nr,
                                                        uint32_t* out)
{
    const int r = blockIdx.x * kBlock + threadIdx.x;
    if (r > nr) return;
    if (dlen && *dlen < len) len = *dlen;
    uint32_t lo = 0, hi = len;
    while (lo < hi)
    {
        const uint32_t mid = (lo + hi) >> 1;
        if ((uint64_t)keys[mid] < (uint64_t)r)
            lo = mid + 1;
        else
            hi = mid;
    }
    out[r] = lo;
}

// triangles grouped by rank, with set-local indices.  Records below 3m gave
// their triangles in rank order already (records of a set lie in its points'
// record range); the spare records' triangles were sorted by rank.
__global__ __launch_bounds__(kBlock) void k_batch_tris(const uint32_t* tris, const uint32_t* dT,
                                                       const uint32_t* dtmain, const uint32_t* trank,
                                                       const uint64_t* skey, const uint32_t* sidx,
                                                       const uint32_t* mstart, const uint32_t* sstart,
                                                       const uint32_t* toff, const uint32_t* in_off, uint32_t* out)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t T = *dT, tmain = min(*dtmain, T);
    if (i >= T) return;
    uint32_t src, dst, r;
    if (i < tmain)
    {
        r = trank[i];
        src = i;
        dst = toff[r] + (i - mstart[r]);
    }
    else
    {
        const uint32_t j = i - tmain;
        r = (uint32_t)skey[j];
        src = tmain + sidx[j];
        dst = toff[r] + (mstart[r + 1] - mstart[r]) + (j - sstart[r]);
    }
    if (dst >= T || src >= T) return; // offsets of another batch's counts (a failed replay guard)
    const uint32_t base = in_off[r];
    out[3 * dst] = tris[3 * src] - base;
    out[3 * dst + 1] = tris[3 * src + 1] - base;
    out[3 * dst + 2] = tris[3 * src + 2] - base;
}

// keys of the triangles of spare records (the ones after the first *dtmain):
// their ranks, padded to cb keys with the rank `pad` (past every set)
__global__ __launch_bounds__(kBlock) void k_spare_keys(const uint32_t* trank, const uint32_t* dT,
                                                       const uint32_t* dtmain, uint32_t cb, uint32_t pad,
                                                       uint64_t* key, uint32_t* idx)
{
    const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= cb) return;
    const uint32_t T = *dT, tmain = min(*dtmain, T);
    key[j] = j < T - tmain ? trank[tmain + j] : pad;
    idx[j] = j;
}

// first output triangle of every rank: the triangles of its records below
// the spare region and those of spare records (one block)
__global__ __launch_bounds__(1024) void k_tri_offsets(const uint32_t* mstart, const uint32_t* sstart, int nr,
                                                      uint32_t* toff)
{
    __shared__ uint32_t carry;
    __shared__ uint32_t wsum[16];
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < nr; base += 1024)
    {
        const int r = base + (int)threadIdx.x;
        const uint32_t v = r < nr ? (mstart[r + 1] - mstart[r]) + (sstart[r + 1] - sstart[r]) : 0;
        const uint32_t lane = lane_id(), wave = threadIdx.x / 64;
        const uint32_t inc = wave_inclusive_scan(v);
        if (lane == 63) wsum[wave] = inc;
        __syncthreads();
        uint32_t before = 0, total = 0;
        for (int w = 0; w < 16; ++w)
        {
            if ((uint32_t)w < wave) before += wsum[w];
            total += wsum[w];
        }
        const uint32_t c = carry;
        if (r < nr) toff[r] = c + before + inc - v;
        __syncthreads();
        if (threadIdx.x == 0) carry = c + total;
        __syncthreads();
    }
    if (threadIdx.x == 0) toff[nr] = carry;
}

// the scalars a run ends with, in one block for one read-back:
// 0 triangles, 1 err, 2 spare demand, 3 cells redone, 4.. the 16 counters
__global__ void k_result_block(const uint32_t* dT, const int* err, const unsigned* spare, const int* redo,
                               const unsigned long long* counters, uint32_t* out)
{
    if (threadIdx.x != 0) return;
    out[0] = *dT;
    out[1] = (uint32_t)*err;
    out[2] = *spare;
    out[3] = redo ? (uint32_t)*redo : 0u;
    unsigned long long* c = reinterpret_cast<unsigned long long*>(out + 4);
    for (int i = 0; i < 16; ++i) c[i] = counters[i];
}

__global__ void k_guard_spare(const uint32_t* dT, const uint32_t* dtmain, uint32_t cb, unsigned long long* guard,
                              unsigned long long bit)
{
    if (threadIdx.x != 0) return;
    const uint32_t T = *dT, tmain = min(*dtmain, T);
    if (T - tmain > cb) atomicOr(guard, bit);
}

__global__ __launch_bounds__(kBlock) void k_rep_local(uint32_t* rep, uint32_t n, const uint32_t* soff, int k)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) rep[i] -= soff[set_of(soff, k, i)];
}

// (total: the distinct count; when every point is distinct each one is its
// own representative, written in order instead of through the sort's
// permutation)
__global__ __launch_bounds__(kBlock) void k_rep(const uint32_t* idx, const uint8_t* F, const uint32_t* dist,
                                                uint32_t n, const uint32_t* ID0, const uint32_t* total,
                                                uint32_t* rep)
{
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= n) return;
    if (*total == n)
    {
        rep[p] = p;
        return;
    }
    const uint32_t d = F[p] ? dist[p] : dist[p] - 1;
    rep[idx[p]] = ID0[d];
}

__global__ __launch_bounds__(kBlock) void k_frame1_keys(const double2* P0, uint32_t m, uint64_t* keys,
                                                        uint32_t* vals)
{
    const uint32_t q = blockIdx.x * kBlock + threadIdx.x;
    if (q >= m) return;
    const uint32_t d = m - 1 - q;
    keys[q] = order_key(P0[d].y);
    vals[q] = d;
}

__global__ __launch_bounds__(kBlock) void k_iota(uint32_t* a, uint32_t n)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) a[i] = i;
}

__global__ __launch_bounds__(kBlock) void k_fill_u32(uint32_t* a, uint32_t n, uint32_t v)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) a[i] = v;
}

// Clears device memory with a kernel: a run can then be captured in a HIP
// graph and replayed (memset nodes are avoided in captured graphs).
__global__ __launch_bounds__(kBlock) void k_fill_words(uint32_t* a, size_t n, uint32_t v)
{
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) a[i] = v;
}

inline Status fill_words(void* p, size_t words, uint32_t v, hipStream_t s)
{
    if (!words) return Status::ok;
    k_fill_words<<<std::min<size_t>(cdiv(words, kBlock), 2048), kBlock, 0, s>>>((uint32_t*)p, words, v);
    return Status::ok;
}

// ---------------------------------------------------------------------------
// 5: kd levels
// ---------------------------------------------------------------------------

// The kd forest.  Level d holds the children (2j, 2j + 1) of level d-1's
// nodes, followed by the roots of the point sets whose trees start at level
// d (a single set: one root at level 0).  The offsets buffer starts with a
// level table, int4 per level: node count, offset of the level's node
// offsets (count + 1 of them) after the table, index of its first node.
constexpr int kMaxLevels = 64;
constexpr int kLvlHdr = 4 * kMaxLevels;
__device__ __forceinline__ uint32_t lcount(const int* off, int d) { return (uint32_t)off[4 * d]; }
__device__ __forceinline__ size_t off_base(const int* off, int d) { return kLvlHdr + (size_t)off[4 * d + 1]; }
__device__ __forceinline__ size_t node_index(const int* off, int d, uint32_t j) { return (size_t)off[4 * d + 2] + j; }
__device__ __forceinline__ int left_size(int s) { return s >= 4 ? s / 2 : s; }

// node offsets of the roots at level d: the point sets of ranks [ra, rb),
// whose distinct points start at vstart[rank]; they follow the children of
// level d-1 (or start the level)
__global__ __launch_bounds__(kBlock) void k_roots(int* off, int d, int ra, int rb, const uint32_t* vstart)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    const int nr = rb - ra;
    if (i > nr) return;
    int* od = off + off_base(off, d);
    const uint32_t first = d > 0 ? 2 * lcount(off, d - 1) : 0;
    od[first + i] = (int)vstart[ra + i]; // i = 0: where the level's children end
}

__global__ __launch_bounds__(kBlock) void k_children(int* off, int d)
{
    const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t cnt = lcount(off, d);
    if (j >= cnt) return;
    const int* od = off + off_base(off, d);
    int* oc = off + off_base(off, d + 1);
    const int o = od[j];
    const int s = od[j + 1] - o;
    oc[2 * j] = o;
    oc[2 * j + 1] = o + left_size(s);
    if (j == cnt - 1) oc[2 * cnt] = od[cnt];
}

// Each kd array carries, beside its vertex ids, their frame-1 ranks (R):
// an id is the point's frame-0 rank, so a point's key in either frame is at
// hand without a lookup.  Within a node both arrays are sorted by their own
// frame's key, so the positional cut of the frame-f array ("the first
// left_size(s) go left") is "key below the key at position o + left_size(s)"
// (medk, per node), which decides the side of every entry of the other array
// from that entry alone.
__global__ __launch_bounds__(kBlock) void k_kd_init(uint32_t* B0, const uint32_t* B1, uint32_t m, uint32_t* R0,
                                                    uint32_t* R1)
{
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= m) return;
    B0[p] = p;
    R1[p] = p;
    R0[B1[p]] = p;
}

struct LoadU8
{
    const uint8_t* F;
    __device__ uint32_t operator()(uint32_t p) const { return F[p]; }
};

// One kd level in one pass: the stable two-way partition, within each node,
// of the array sorted in the other frame (position p goes left when its key,
// the id on levels cut in frame 0 and the frame-1 rank on the others, is
// below its node's medk), ids and frame-1 ranks together, with the count of
// "goes left" before p in its node from a segmented scan whose segments are
// the level's nodes.  Position p's node is the last j with od[j] <= p (the
// level's nodes cover [0, od[cnt]) in order; positions from od[cnt] on
// belong to sets rooted deeper and keep their place), so no node array is
// read or written.
// A tile of kKdTile positions: each wave owns a contiguous run of
// 64 kKdItems positions and takes it in 64-position steps (lane l holds
// position run + 64 k + l of step k), so loads are whole lines and the
// stores of a step fall into one run per side of each node it touches.
// The tile's count since its first position's node start comes from the
// tiles before it by a decoupled look-back (tiles numbered in start order by
// an atomic ticket, as in the radix passes): a tile publishes the count
// since its last node start (flagged when it holds one) and, once it knows
// its carry, its inclusive value; a tile reads the words of the 64 tiles
// before it at once, and the nearest that is inclusive or holds a node start
// ends the look-back (the counts up to it add up).
// state: this level's look-back words (a ticket, 63 pad words, one word per
// tile; cleared); next: the other level's, cleared here for the next level.
// A look-back that polls one word kLbSpins times sets kKdErr and goes on.
constexpr int kKdItems = 16;
constexpr int kKdTile = kBlock * kKdItems;
constexpr int kKdNodes = 512;              // node offsets of a tile kept in LDS (beyond: read from HBM)
constexpr uint32_t kKdErr = 1u << 24;      // err bit: a kd level's look-back gave up
constexpr uint32_t kKdHead = 1u << 29;     // segmented value: a node starts inside (count since its start)
constexpr uint32_t kKdCount = kKdHead - 1;

__host__ __device__ inline size_t kd_half_words(size_t m) { return 64 + (m + kKdTile - 1) / kKdTile; }

// segmented "then": b after a (a's count is dropped when b holds a node start)
__device__ __forceinline__ uint32_t kd_seg(uint32_t a, uint32_t b) { return (b & kKdHead) ? b : a + b; }

__global__ __launch_bounds__(kBlock) void k_kd_level(const uint32_t* Bo, const uint32_t* Ro, const int* off, int d,
                                                     const uint32_t* medk, uint32_t m, uint32_t* Bout,
                                                     uint32_t* Rout, uint32_t* state, uint32_t* next, int* err)
{
    __shared__ uint32_t s_tile, s_carry;
    __shared__ int s_j0, s_nl;
    __shared__ int s_od[kKdNodes + 1];
    __shared__ uint32_t s_mk[kKdNodes];
    __shared__ uint32_t wagg[kWaves];
    const uint32_t t = threadIdx.x, lane = lane_id(), wave = t / 64;
    const int cnt = (int)lcount(off, d);
    const int* od = off + off_base(off, d);
    if (t == 0) s_tile = atomicAdd(state, 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    next[64 + tile] = 0;
    if (tile == 0 && t == 0) next[0] = 0;
    const uint32_t t0 = tile * kKdTile;
    const uint32_t run0 = t0 + wave * (64 * kKdItems) + lane; // step k: run0 + 64 k
    uint32_t v[kKdItems], r[kKdItems];
#pragma unroll
    for (int k = 0; k < kKdItems; ++k)
    {
        const uint32_t p = run0 + 64 * k;
        v[k] = p < m ? Bo[p] : 0;
        r[k] = p < m ? Ro[p] : 0;
    }
    const int end = od[cnt];
    // the node of the tile's first position (wave 0): the nodes of a level
    // have nearly equal sizes, so 64 offsets around the proportional guess
    // hold it; otherwise a 64-way search
    if (wave == 0)
    {
        int lo = cnt - 1;
        if ((int)t0 < end)
        {
            const int g = (int)(((unsigned long long)t0 * (unsigned)cnt) / (unsigned)end);
            const int a = max(0, min(cnt - 1, g) - 31);
            const int idx = a + (int)lane;
            const uint64_t le = __ballot(idx <= cnt && od[idx] <= (int)t0);
            if ((le & 1) && ~le)
                lo = a + __popcll(le) - 1;
            else
            {
                lo = 0;
                int hi = cnt; // od[lo] <= t0 < od[hi] (od[0] = 0)
                while (hi - lo > 1)
                {
                    const int step = (hi - lo + 63) / 64;
                    const int ix = lo + (int)lane * step;
                    const uint64_t bl = __ballot(ix < hi && od[ix] <= (int)t0);
                    const int c = max(__popcll(bl), 1);
                    lo += (c - 1) * step;
                    hi = min(hi, lo + step);
                }
            }
        }
        if (lane == 0)
        {
            s_j0 = lo;
            s_nl = min(kKdNodes, cnt - lo);
        }
    }
    __syncthreads();
    const int j0 = s_j0, nl = s_nl;
    for (int i = (int)t; i <= nl; i += kBlock)
    {
        s_od[i] = od[j0 + i];
        if (i < nl) s_mk[i] = medk[j0 + i];
    }
    __syncthreads();
    // offsets and keys of nodes beyond the cached ones come from HBM
    auto odv = [&](int j) { return j - j0 <= nl ? s_od[j - j0] : od[j]; };
    auto mkv = [&](int j) { return j - j0 < nl ? s_mk[j - j0] : medk[j]; };
    // this lane's first node: last j with od[j] <= run0
    int j = j0;
    if ((int)run0 < end)
    {
        int lo, hi;
        if ((int)run0 >= s_od[nl])
        {
            lo = j0 + nl;
            hi = cnt - 1;
        }
        else
        {
            lo = j0;
            hi = j0 + nl - 1;
        }
        while (lo < hi)
        {
            const int mid = (lo + hi + 1) >> 1;
            if (odv(mid) <= (int)run0)
                lo = mid;
            else
                hi = mid - 1;
        }
        j = lo;
    }
    // each step's node, side and node start; the wave's run: the count of
    // lefts after its last node start (flagged when it holds one)
    const bool odd = d & 1;
    int nj[kKdItems];
    uint32_t flags = 0, heads = 0;
    {
        int e = odv(j + 1);
        uint32_t mk = mkv(j);
        int o = odv(j);
        int last_head = -1; // position of the run's last node start
#pragma unroll
        for (int k = 0; k < kKdItems; ++k)
        {
            const int p = (int)run0 + 64 * k;
            nj[k] = -1; // passes through
            if (p >= end) continue;
            while (p >= e && j + 1 < cnt)
            {
                ++j;
                o = e;
                e = odv(j + 1);
                mk = mkv(j);
            }
            nj[k] = j;
            if ((odd ? r[k] : v[k]) < mk) flags |= 1u << k;
            if (p == o)
            {
                heads |= 1u << k;
                last_head = p;
            }
        }
        // the wave's last node start, then its lefts from there on
        int lh = last_head;
#pragma unroll
        for (int dd = 32; dd >= 1; dd >>= 1) lh = max(lh, __shfl_xor(lh, dd, 64));
        uint32_t c = 0;
#pragma unroll
        for (int k = 0; k < kKdItems; ++k)
            c += ((flags >> k) & 1) && (int)run0 + 64 * k >= lh ? 1u : 0u;
#pragma unroll
        for (int dd = 32; dd >= 1; dd >>= 1) c += __shfl_xor(c, dd, 64);
        if (lane == 0) wagg[wave] = (lh >= 0 ? kKdHead : 0u) | c;
    }
    __syncthreads();
    uint32_t before = 0, total = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w)
    {
        if ((uint32_t)w < wave) before = kd_seg(before, wagg[w]);
        total = kd_seg(total, wagg[w]);
    }
    // the tile's carry (wave 0)
    if (wave == 0)
    {
        uint32_t* const st = state + 64;
        uint32_t carry = 0;
        if (tile == 0)
        {
            if (lane == 0)
                __hip_atomic_store(&st[0], kLbInc | (total & kKdCount), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        else
        {
            if (lane == 0) __hip_atomic_store(&st[tile], kLbAgg | total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            int base = (int)tile - 1;
            uint32_t spins = 0;
            for (;;)
            {
                const int jt = base - (int)lane;
                const uint32_t w = jt >= 0 ? __hip_atomic_load(&st[jt], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                           : kLbInc; // before tile 0: nothing
                const bool ready = (w & (kLbAgg | kLbInc)) != 0;
                const uint64_t stops = __ballot(ready && (w & (kLbInc | kKdHead)));
                const uint64_t waiting = __ballot(!ready);
                const int sl = stops ? __ffsll((unsigned long long)stops) - 1 : 64;
                const uint64_t need = sl >= 63 ? ~0ull : ((2ull << sl) - 1);
                if (waiting & need)
                {
                    if (++spins >= kLbSpins)
                    {
                        if (lane == 0) atomicOr(err, (int)kKdErr);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                uint32_t c = (int)lane <= sl ? (w & kKdCount) : 0;
#pragma unroll
                for (int dd = 32; dd >= 1; dd >>= 1) c += __shfl_xor(c, dd, 64);
                carry += c;
                if (sl < 64) break;
                base -= 64;
            }
            if (lane == 0)
                __hip_atomic_store(&st[tile], kLbInc | (kd_seg(carry, total) & kKdCount), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        }
        if (lane == 0) s_carry = carry;
    }
    __syncthreads();
    // each step: a segmented scan across the wave gives every position the
    // count of lefts before it in its node; the destinations follow
    uint32_t carry = kd_seg(s_carry, before) & kKdCount;
    const uint64_t lt = lanemask_lt();
#pragma unroll
    for (int k = 0; k < kKdItems; ++k)
    {
        const uint32_t p = run0 + 64 * k;
        const bool left = (flags >> k) & 1;
        const uint64_t hb = __ballot((heads >> k) & 1);
        const uint64_t lb = __ballot(left);
        // lefts before this lane since the nearest node start at or before it
        const uint64_t hb_le = hb & (lt | (1ull << lane));
        uint32_t z;
        if (hb_le)
        {
            const int h = 63 - __clzll((long long)hb_le); // the lane of that node start
            z = (uint32_t)__popcll(lb & lt & ~((1ull << h) - 1));
        }
        else
            z = carry + (uint32_t)__popcll(lb & lt);
        // the next step's carry: the count after the step's last node start
        if (hb)
        {
            const int h = 63 - __clzll((long long)hb);
            carry = (uint32_t)__popcll(lb & ~((1ull << h) - 1));
        }
        else
            carry += (uint32_t)__popcll(lb);
        if (p >= m) continue;
        uint32_t dst = p;
        if (nj[k] >= 0)
        {
            const int o = odv(nj[k]);
            const int ls = left_size(odv(nj[k] + 1) - o);
            dst = left ? (uint32_t)(o + (int)z) : (uint32_t)(o + ls + ((int)p - o - (int)z));
        }
        if (dst < m)
        {
            Bout[dst] = v[k];
            Rout[dst] = r[k];
        }
    }
}

// Vertex ids renumbered into the order of the frame-0 array at the cell
// level: every kd node's points get a contiguous range of ids (their
// positions), so the coordinates that the cells, merges, anchor walks and
// seam pieces read lie together in memory (as x-ranks, a node's points are
// spread over the whole width of its slab).  newid[v] = p for v = B0[p];
// coordinates and original indices are written in the new order (into the
// spare buffers, which then swap roles with the old ones) and B0 becomes the
// identity.  B1 keeps the old ids: its readers map them through newid.
__global__ __launch_bounds__(kBlock) void k_renum(uint32_t* B0, uint32_t m, const double2* P0, const uint32_t* ID0,
                                                  uint32_t* newid, double2* P0n, uint32_t* ID0n)
{
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= m) return;
    const uint32_t v = B0[p];
    newid[v] = p;
    P0n[p] = P0[v];
    ID0n[p] = ID0[v];
    B0[p] = p;
}

// vertex ids in a[0, na) mapped (entries that are not ids are left alone)
__global__ __launch_bounds__(kBlock) void k_renum_ids(const uint32_t* newid, uint32_t m, int* a, size_t na)
{
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < na; i += (size_t)gridDim.x * kBlock)
    {
        const int v = a[i];
        if ((unsigned)v < m) a[i] = (int)newid[v];
    }
}

// ---------------------------------------------------------------------------
// 6-7: leaves and merges
// ---------------------------------------------------------------------------

using GQE = QuadEdges<int*, const double2*>;

// A merged node's result as its parent needs it.  k_seg_final stores
// (-(ldo + 2), rdo, pool) when the hull edges still have to be re-walked to
// the parent's frame; the re-walk (builder.hxx:340-346) is then done by the
// first reader (the anchor kernel's tangent lane or k_merge), whose launch
// hides its latency, and written back.  c: the child's index at level dc.
__device__ __forceinline__ int4 child_result(GQE& q, int4* res, const int2* ext, const int* off, int dc, uint32_t c)
{
    const size_t ni = node_index(off, dc, c);
    int4 r = res[ni];
    if (r.x <= -2)
    {
        const int* od = off + off_base(off, dc);
        const int sz = od[c + 1] - od[c];
        const int2 x = ext[ni];
        int first, second;
        q.rewalk_ext(-r.x - 2, r.y, dc & 1, (dc - 1) & 1, sz, x.x, x.y, first, second);
        r.x = first;
        r.y = second;
        res[ni] = r;
    }
    return r;
}

__global__ __launch_bounds__(kMergeBlock) void k_leaves(int* NX, int* OR, const double2* PT, int* VE,
                                                        const uint32_t* B0, const uint32_t* B1, const int* off, int D,
                                                        uint32_t m, int lim_e, int4* res, int* err,
                                                        unsigned long long* counters)
{
    const uint32_t j = blockIdx.x * kMergeBlock + threadIdx.x;
    GQE q;
    q.init(NX, OR, PT, VE, lim_e, (int)m);
    if (j < lcount(off, D))
    {
        const int* oD = off + off_base(off, D);
        const int o = oD[j];
        const int s = oD[j + 1] - o;
        if (s >= 1)
        {
            int db = D;
            uint32_t jb = j;
            while (db > 0 && jb < 2 * lcount(off, db - 1)) // roots have no parent
            {
                const uint32_t pj = jb >> 1;
                const int* op = off + off_base(off, db - 1);
                if (op[pj + 1] - op[pj] >= 4) break;
                --db;
                jb = pj;
            }
            const int out_axis = db > 0 ? ((db - 1) & 1) : 0;
            const uint32_t* B = out_axis ? B1 : B0;
            int v[3];
            for (int i = 0; i < s; ++i) v[i] = (int)B[o + i];
            int first, second;
            Pool pool;
            q.leaf(v, s, 3 * o, first, second, pool);
            res[node_index(off, db, jb)] = make_int4(first, second, pool.head, pool.tail);
        }
    }
    flush_counts(counters, q.pc);
    if (q.err) atomicOr(err, 2 << q.err);
}

// One merge per lane on the first kMergeLanes lanes of each wavefront: the
// merges are diverging chains of dependent loads, and a level has only a few
// thousand of them, so fewer per wavefront spread them over more SIMDs (16,
// 32 or 64 per wavefront, tuned per workload; 8 or 4 on levels too small to
// give each SIMD two wavefronts).
constexpr uint32_t kMergeWavesMin = 2048;
template <uint32_t kMergeLanes>
__global__ __launch_bounds__(kMergeBlock) void k_merge(int* NX, int* OR, const double2* PT, int* VE, const int* off,
                                                       int d, uint32_t m, int lim_e, int4* res, const int2* ext,
                                                       int* err, unsigned long long* counters)
{
    const uint32_t j = blockIdx.x * kMergeLanes + threadIdx.x;
    GQE q;
    q.init(NX, OR, PT, VE, lim_e, (int)m);
    if (threadIdx.x < kMergeLanes && j < lcount(off, d))
    {
        const int* od = off + off_base(off, d);
        const int s = od[j + 1] - od[j];
        if (s >= 4)
        {
            const int4 L = child_result(q, res, ext, off, d + 1, 2 * j);
            const int4 R = child_result(q, res, ext, off, d + 1, 2 * j + 1);
            Pool pool = q.cat(Pool{ L.z, L.w }, Pool{ R.z, R.w });
            const int axis = d & 1;
            const int out_axis = d > 0 ? ((d - 1) & 1) : 0;
            const int2 x = d > 0 ? ext[node_index(off, d, j)] : make_int2(-1, -1);
            int first, second;
            q.merge_node(L.x, L.y, R.x, R.y, axis, out_axis, s, first, second, pool, x.x, x.y);
            res[node_index(off, d, j)] = make_int4(first, second, pool.head, pool.tail);
        }
    }
    flush_counts(counters, q.pc);
    flush_connects(counters, q.pc);
    if (q.err) atomicOr(err, 2 << q.err);
}


// ---------------------------------------------------------------------------
// 6-7 in LDS: cells.  One wavefront builds a whole kd subtree of at most
// kCellMax points: its kd partition, leaves and merges run on a copy of the
// cell's points and quad-edge records in LDS, and the records are written to
// HBM once.  The steps are the same as k_leaves / k_merge (same cuts, same
// leaf and merge code, same record ranges), so the store that leaves the
// cell is the one the level kernels would have built.
// ---------------------------------------------------------------------------

constexpr int kCellMax = 512;             // points per cell at most (the tuned size is 128, 256 or 512)
constexpr int kCellLevels = 12;

// The fast pass runs the filters alone: no exact stage, no calls, fewer
// registers; a cell where a filter cannot decide is listed and rebuilt by
// the exact pass.
using LQE = QuadEdges<short*, const double2*, false, false, true>; // cell records in LDS: local ids fit 16 bits
using LQX = QuadEdges<short*, const double2*, false, true>;

// cell c -> (level, node index): the nodes of level dc, then the roots of
// the sets rooted below dc
__device__ __forceinline__ bool cell_node(const int* off, int dc, int D, uint32_t c, int& dl, uint32_t& j)
{
    const uint32_t n0 = lcount(off, dc);
    if (c < n0)
    {
        dl = dc;
        j = c;
        return true;
    }
    if (dc >= D) return false;
    c -= n0;
    for (int d = dc + 1; d <= D; ++d)
    {
        const uint32_t first = 2 * lcount(off, d - 1);
        const uint32_t roots = lcount(off, d) - first;
        if (c < roots)
        {
            dl = d;
            j = first + c;
            return true;
        }
        c -= roots;
    }
    return false;
}

// Node slots of a cell's sub-levels (heap order): the partition of s points
// stops below the level whose nodes all hold fewer than four, so the deepest
// level L has 2^L < 2s/3 nodes and levels 0..L take 2^(L+1) - 1 < 4s/3 slots
// (checked for every s <= cap <= 512 on the host; the partition loop still
// checks it).
__host__ __device__ constexpr int cell_nodes(int cap) { return (4 * cap + 2) / 3 + 2; }

// LDS of a cell of at most cap points (bytes), laid out by CellLds
__host__ __device__ constexpr size_t cell_lds_bytes(int cap)
{
    return ((16 * (size_t)cap + 24 * (size_t)cap + 12 * (size_t)cap + 4 * (size_t)cap +
             2 * (size_t)(cell_nodes(cap) + kCellLevels + 2) + 7) & ~(size_t)7) +
        8 * (size_t)cell_nodes(cap) + 64;
}

struct CellLds
{
    double2* P;   // cap points, frame-0 order
    short* LNX;   // 3 cap records, onext of the 4 directed edges each
    short* LOR;   // origins of the 2 primal edges of each record
    short *A0, *A1; // the cell's positions in frame-0 / frame-1 order (after the
                    // merges: an int per vertex, its lowest out-going edge)
    short* loff;  // sub-level node offsets
    short4* lres; // sub-level node results (cell_nodes(cap) slots)
    // partition scratch, inside the record area (unused until the leaves)
    short *T, *Z;
    uint8_t* right;
    uint32_t* rk; // frame-1 ranks of the cell's points (load phase)
    __device__ CellLds(unsigned char* lds, int cap)
    {
        P = (double2*)lds;
        LNX = (short*)(lds + 16 * (size_t)cap);
        LOR = LNX + 12 * cap;
        A0 = LOR + 6 * cap;
        A1 = A0 + cap;
        loff = A1 + cap;
        lres = (short4*)(lds + ((16 * (size_t)cap + 24 * (size_t)cap + 12 * (size_t)cap + 4 * (size_t)cap +
                                 2 * (size_t)(cell_nodes(cap) + kCellLevels + 2) + 7) & ~(size_t)7));
        unsigned char* x = (unsigned char*)LNX; // 24 cap bytes
        T = (short*)x;
        Z = (short*)(x + 2 * (size_t)cap);
        right = x + 4 * (size_t)cap + 4;
        rk = (uint32_t*)(x + 8 * (size_t)cap);
    }
};

// One cell: returns false (and writes nothing) when Q lacks the exact stage
// and some filter could not decide.
// kCellPer: positions per lane in the cell's scans (cap <= 64 kCellPer)
// B0: the frame-0 array (vertex ids, which after the renumbering are its
// positions); rank1: beside it, each entry's frame-1 rank.
// Inlined into each kernel that calls it (the cell kernels and the small-set
// kernel): an out-of-line copy shared by two kernels takes the larger
// kernel's register budget (k_cell 110 -> 248 VGPRs, 6.0 -> 9.1 ms).
template <class Q, int kCellPer>
__device__ __forceinline__ bool cell_build(unsigned char* lds, int cap, uint32_t cell, int* NX, int* OR,
                                           const double2* PT, int* VE,
                           const uint32_t* B0, const uint32_t* rank1, const int* off, int dc, int D, int4* res,
                           int* err, unsigned long long* counters)
{
    __shared__ int lbase[kCellLevels + 1]; // first offset of each sub-level in loff
    __shared__ int bad;
    const CellLds W(lds, cap);
    double2* P = W.P;
    short *A0 = W.A0, *A1 = W.A1, *T = W.T, *Z = W.Z, *loff = W.loff, *LNX = W.LNX, *LOR = W.LOR;
    uint8_t* right = W.right;
    uint32_t* rk = W.rk;
    short4* lres = W.lres;
    const int kNodes = cell_nodes(cap);

    const int lane = threadIdx.x;
    int dl;
    uint32_t j;
    if (!cell_node(off, dc, D, cell, dl, j)) return true;
    const int* od = off + off_base(off, dl);
    const int o = od[j];
    const int s = od[j + 1] - o;
    if (s <= 0) return true;
    if (s > cap)
    {
        if (lane == 0) atomicOr(err, 2 << 1);
        return true;
    }
    // points in frame-0 order; vertex ids within a node ascend in that order
    for (int i = lane; i < s; i += 64)
    {
        const int v = (int)B0[o + i];
        P[i] = PT[v];
        A0[i] = (short)i;
        rk[i] = rank1[o + i];
    }
    if (lane == 0)
    {
        bad = 0;
        lbase[0] = 0;
        loff[0] = 0;
        loff[1] = (short)s;
    }
    __syncthreads();
    // frame 1: the cell's points ordered by their frame-1 ranks (distinct),
    // each placed at the count of smaller ranks in the cell (read from LDS,
    // no gather from HBM)
    {
        uint32_t kk[kCellPer];
        int rr[kCellPer];
#pragma unroll
        for (int r = 0; r < kCellPer; ++r)
        {
            const int i = lane + 64 * r;
            kk[r] = i < s ? rk[i] : 0xffffffffu;
            rr[r] = 0;
        }
        for (int jx = 0; jx < s; ++jx)
        {
            const uint32_t kj = rk[jx];
#pragma unroll
            for (int r = 0; r < kCellPer; ++r) rr[r] += kj < kk[r] ? 1 : 0;
        }
#pragma unroll
        for (int r = 0; r < kCellPer; ++r)
        {
            const int i = lane + 64 * r;
            if (i < s) A1[rr[r]] = (short)i;
        }
    }
    // the cell's extent: every coordinate difference of two of its points is
    // at most wx in x and wy in y, which fixes the filters' bounds for the cell
    double wx = 0, wy = 0;
    if constexpr (!Q::kHasExact)
    {
        double x0 = INFINITY, x1 = -INFINITY, y0 = INFINITY, y1 = -INFINITY;
        for (int i = lane; i < s; i += 64)
        {
            const double2 p = P[i];
            x0 = fmin(x0, p.x);
            x1 = fmax(x1, p.x);
            y0 = fmin(y0, p.y);
            y1 = fmax(y1, p.y);
        }
#pragma unroll
        for (int dd = 32; dd >= 1; dd >>= 1)
        {
            x0 = fmin(x0, __shfl_xor(x0, dd, 64));
            x1 = fmax(x1, __shfl_xor(x1, dd, 64));
            y0 = fmin(y0, __shfl_xor(y0, dd, 64));
            y1 = fmax(y1, __shfl_xor(y1, dd, 64));
        }
        wx = x1 - x0;
        wy = y1 - y0;
    }
    __syncthreads();
    // kd partition of the cell, sub-level by sub-level (absolute level dl + l)
    int L = 0;
    for (;;)
    {
        const int cnt = 1 << L;
        const short* ol = loff + lbase[L];
        bool mine = false;
        for (int q = lane; q < cnt; q += 64) mine |= ol[q + 1] - ol[q] >= 4;
        if (__ballot(mine) == 0) break;
        if (L + 1 >= kCellLevels || (4 << L) - 1 > kNodes || lbase[L] + 3 * cnt + 2 > kNodes + kCellLevels)
        {
            bad = 1; // deeper than the slots allow: cannot happen for cap points
            break;
        }
        const int axis = (dl + L) & 1;
        short* Af = axis ? A1 : A0;
        short* Ao = axis ? A0 : A1;
        // children offsets
        const int nb = lbase[L] + cnt + 1;
        short* oc = loff + nb;
        for (int qn = lane; qn < cnt; qn += 64)
        {
            const int a = ol[qn], sz = ol[qn + 1] - a;
            oc[2 * qn] = (short)a;
            oc[2 * qn + 1] = (short)(a + left_size(sz));
        }
        if (lane == 0)
        {
            lbase[L + 1] = nb;
            oc[2 * cnt] = ol[cnt];
        }
        __syncthreads();
        // membership from the positional split of the frame array
        for (int i = lane; i < s; i += 64)
        {
            int lo = 0, hi = cnt - 1; // node of position i: last q with ol[q] <= i
            while (lo < hi)
            {
                const int mid = (lo + hi + 1) >> 1;
                if (ol[mid] <= i)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            const int a = ol[lo], sz = ol[lo + 1] - a;
            right[Af[i]] = (uint8_t)((i - a) >= left_size(sz) ? 1 : 0);
        }
        __syncthreads();
        // stable partition of the other array: exclusive count of lefts
        // (kCellPer consecutive positions per lane, one wave-wide scan)
        {
            int f[kCellPer];
            int x = 0;
#pragma unroll
            for (int r = 0; r < kCellPer; ++r)
            {
                const int i = kCellPer * lane + r;
                f[r] = i < s ? (right[Ao[i]] ? 0 : 1) : 0;
                x += f[r];
            }
            const int own = x;
#pragma unroll
            for (int dd = 1; dd < 64; dd <<= 1)
            {
                const int y = __shfl_up(x, dd, 64);
                if (lane >= dd) x += y;
            }
            int ex = x - own;
#pragma unroll
            for (int r = 0; r < kCellPer; ++r)
            {
                const int i = kCellPer * lane + r;
                if (i < s) Z[i] = (short)ex;
                ex += f[r];
            }
            if (lane == 63) Z[s] = (short)x;
        }
        __syncthreads();
        for (int i = lane; i < s; i += 64)
        {
            int lo = 0, hi = cnt - 1;
            while (lo < hi)
            {
                const int mid = (lo + hi + 1) >> 1;
                if (ol[mid] <= i)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            const int a = ol[lo], sz = ol[lo + 1] - a;
            const int v = Ao[i];
            const int z = Z[i] - Z[a];
            T[right[v] ? a + left_size(sz) + (i - a - z) : a + z] = (short)v;
        }
        __syncthreads();
        for (int i = lane; i < s; i += 64) Ao[i] = T[i];
        __syncthreads();
        ++L;
    }
    // A cell of at most three points is a leaf whose result belongs to its
    // highest ancestor still below four points (the node its parent merges,
    // as in k_leaves); roots have no parent.
    int tdl = dl;
    uint32_t tj = j;
    if (s < 4)
        while (tdl > 0 && tj < 2 * lcount(off, tdl - 1))
        {
            const int* op = off + off_base(off, tdl - 1);
            const uint32_t pj = tj >> 1;
            if (op[pj + 1] - op[pj] >= 4) break;
            --tdl;
            tj = pj;
        }
    Q q;
    // no vertex-edge upkeep inside the cell: the entries are derived from the
    // final records below
    q.init(LNX, LOR, P, nullptr, 4 * 3 * s, s);
    if constexpr (!Q::kHasExact)
        if (wx > 0 && wy > 0 && wx < INFINITY && wy < INFINITY) static_bounds(wx, wy, q.sb_orient, q.sb_incircle);
    // without usable static bounds (an extent of zero, or products beyond the
    // filters' range) the cell goes to the exact rebuild
    if constexpr (Q::kStaticBounds)
        if (!(q.sb_orient > 0 && q.sb_incircle > 0)) return false;
    // leaves at sub-level L (absolute dl + L); the partition scratch that
    // shares the records' space is dead by now (the loop above ends with a
    // barrier after its last use)
    {
        const int cnt = 1 << L;
        const short* ol = loff + lbase[L];
        for (int jl = lane; jl < cnt; jl += 64)
        {
            const int a = ol[jl], sz = ol[jl + 1] - a;
            if (sz < 1) continue;
            int db = L; // highest ancestor (sub-level) still below four points
            int jb = jl;
            while (db > 0)
            {
                const short* op = loff + lbase[db - 1];
                if (op[(jb >> 1) + 1] - op[jb >> 1] >= 4) break;
                --db;
                jb >>= 1;
            }
            const int abs_db = db > 0 ? dl + db : tdl;
            const int out_axis = abs_db > 0 ? ((abs_db - 1) & 1) : 0;
            const short* B = out_axis ? A1 : A0;
            int v[3] = { 0, 0, 0 };
            for (int i = 0; i < sz && i < 3; ++i) v[i] = B[a + i];
            int first, second;
            Pool pool;
            q.leaf(v, sz, 3 * a, first, second, pool);
            lres[(1 << db) - 1 + jb] = make_short4((short)first, (short)second, (short)pool.head, (short)pool.tail);
        }
    }
    __syncthreads();
    // merges, deepest sub-level first
    for (int l = L - 1; l >= 0; --l)
    {
        const int cnt = 1 << l;
        const short* ol = loff + lbase[l];
        for (int jl = lane; jl < cnt; jl += 64)
        {
            const int sz = ol[jl + 1] - ol[jl];
            if (sz < 4) continue;
            const short4 Lr = lres[(2 << l) - 1 + 2 * jl];
            const short4 Rr = lres[(2 << l) - 1 + 2 * jl + 1];
            Pool pool = q.cat(Pool{ Lr.z, Lr.w }, Pool{ Rr.z, Rr.w });
            const int abs_d = dl + l;
            const int axis = abs_d & 1;
            const int out_axis = abs_d > 0 ? ((abs_d - 1) & 1) : 0;
            int first, second;
            q.merge_node(Lr.x, Lr.y, Rr.x, Rr.y, axis, out_axis, sz, first, second, pool);
            lres[(1 << l) - 1 + jl] = make_short4((short)first, (short)second, (short)pool.head, (short)pool.tail);
        }
        __syncthreads();
    }
    // a filter that could not decide: nothing is written, the caller lists the cell
    if (__ballot(q.err == 5) != 0) return false;
    // an edge out of every vertex, the lowest id among its edges (so the
    // choice does not depend on lane timing), in the frame orders' space
    // (A0, A1: 2 cap shorts, dead since the leaves)
    int* ve = (int*)A0;
    for (int i = lane; i < s; i += 64) ve[i] = INT_MAX;
    __syncthreads();
    for (int k = lane; k < 3 * s; k += 64)
    {
        const int o0 = LOR[2 * k];
        if (o0 < 0) continue;
        atomicMin(&ve[o0], 4 * k);
        atomicMin(&ve[LOR[2 * k + 1]], 4 * k + 2);
    }
    __syncthreads();
    // write the records, vertex edges and the cell's result
    const int E0 = 12 * o; // local edge e is global E0 + e
    const int R0 = 3 * o;  // local record k is global R0 + k
    for (int k = lane; k < 3 * s; k += 64)
    {
        const int g = R0 + k;
        const int o0 = LOR[2 * k], o1 = LOR[2 * k + 1];
        if (o0 < 0)
        {
            const int nx = LNX[4 * k];
            NX[4 * g] = nx >= 0 ? R0 + nx : -1;
            OR[2 * g] = -1;
            OR[2 * g + 1] = -1;
            continue;
        }
        for (int r = 0; r < 4; ++r) NX[4 * g + r] = E0 + LNX[4 * k + r];
        OR[2 * g] = (int)B0[o + o0];
        OR[2 * g + 1] = (int)B0[o + o1];
    }
    for (int i = lane; i < s; i += 64) VE[B0[o + i]] = ve[i] != INT_MAX ? E0 + ve[i] : -1;
    if (lane == 0)
    {
        const short4 r = lres[0];
        res[node_index(off, tdl, tj)] = make_int4(r.x >= 0 ? E0 + r.x : -1, r.y >= 0 ? E0 + r.y : -1,
                                                r.z >= 0 ? R0 + r.z : -1, r.w >= 0 ? R0 + r.w : -1);
        if (bad) q.err = q.err ? q.err : 1;
    }
    flush_counts(counters, q.pc);
    if (q.err) atomicOr(err, 2 << q.err);
    return true;
}

// every cell, filters only; undecided cells go to redo (count, then ids)
template <int kCellPer>
__global__ __launch_bounds__(64) void k_cell(int* NX, int* OR, const double2* PT, int* VE, const uint32_t* B0,
                                             const uint32_t* rank1, const int* off, int dc, int D, int cap,
                                             int4* res, int* redo, int* err, unsigned long long* counters)
{
    extern __shared__ __align__(16) unsigned char cell_lds[];
    if (!cell_build<LQE, kCellPer>(cell_lds, cap, blockIdx.x, NX, OR, PT, VE, B0, rank1, off, dc, D, res, err,
                                   counters) &&
        threadIdx.x == 0)
    {
        const int k = atomicAdd(redo, 1);
        redo[1 + k] = (int)blockIdx.x;
    }
}

// the listed cells again, with the exact predicates
template <int kCellPer>
__global__ __launch_bounds__(64) void k_cell_exact(int* NX, int* OR, const double2* PT, int* VE, const uint32_t* B0,
                                                   const uint32_t* rank1, const int* off, int dc, int D, int cap,
                                                   int4* res, const int* redo, int* err,
                                                   unsigned long long* counters)
{
    extern __shared__ __align__(16) unsigned char cell_lds[];
    const int n = redo[0];
    for (int i = blockIdx.x; i < n; i += gridDim.x)
    {
        cell_build<LQX, kCellPer>(cell_lds, cap, (uint32_t)redo[1 + i], NX, OR, PT, VE, B0, rank1, off, dc, D,
                                  res, err, counters);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Small sets: the whole pipeline in one wavefront
// ---------------------------------------------------------------------------
// points of a single set run by k_small at most: from ~300 points on, the
// one wavefront's serial merges take longer than the whole general pipeline
// (measured on MI355X: 256 points 0.50 ms here, 769 points 1.23 ms there)
constexpr size_t kSmallMax = 256;

// A single set of at most 64 kPer points is one cell, so its triangulation
// needs no kd level kernels, no seam pieces and no host decisions: one
// wavefront sorts the points in LDS (x, y, then input index: the loader's
// order, builder.hxx:371-376), keeps the lowest index of each run of equal
// points (builder.hxx:377-388), ranks the distinct points in frame 1 (y
// ascending, x descending, builder.hxx:75), builds the cell as k_cell does
// (filters first, the exact predicates if a filter cannot decide) and
// extracts the faces in record order as the extraction kernels do.  A call
// is then one launch, instead of the general pipeline's dozens of launches
// and host round trips.

// inclusive sum over the wavefront
__device__ __forceinline__ int wave_incl_sum(int v)
{
    const int lane = (int)__lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1)
    {
        const int t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

// out[0] triangles, out[1] distinct points, out[2] flags (1: a coordinate is
// not finite, 2: the cell needed the exact predicates)
template <int kPer>
__global__ __launch_bounds__(64) void k_small(const double2* pts, uint32_t n, int* NX, int* OR, double2* P0, int* VE,
                                              uint32_t* B0, uint32_t* rank1, uint32_t* ID0, uint32_t* rep, int* off,
                                              int4* res, uint32_t* tris, uint32_t tri_cap, uint32_t* out, int* err,
                                              unsigned long long* counters)
{
    constexpr int kCap = 64 * kPer;
    extern __shared__ __align__(16) unsigned char cell_lds[];
    uint64_t* kx = reinterpret_cast<uint64_t*>(cell_lds);
    uint64_t* ky = kx + kCap;
    int* ord = reinterpret_cast<int*>(ky + kCap);
    int* dv = ord + kCap; // distinct point -> input index
    __shared__ int s_flags;
    const int lane = (int)threadIdx.x;
    if (lane == 0)
    {
        s_flags = 0;
        *err = 0;
    }
    if (lane < 16) counters[lane] = 0;
    __syncthreads();
    for (int i = lane; i < (int)n; i += 64)
    {
        const double2 p = pts[i];
        if (!isfinite(p.x) || !isfinite(p.y)) atomicOr(&s_flags, 1);
        kx[i] = order_key(p.x);
        ky[i] = order_key(p.y);
    }
    __syncthreads();
    if (s_flags & 1)
    {
        if (lane == 0)
        {
            out[0] = 0;
            out[1] = 0;
            out[2] = 1;
        }
        return;
    }
    // sorted position of every point: the count of points before it in
    // (x, y, index) order (each lane ranks kPer points against all n, read
    // from LDS; for sets this small a count beats a sorting network's
    // rounds of dependent LDS accesses)
    {
        uint64_t ox[kPer], oy[kPer];
        int rk[kPer];
#pragma unroll
        for (int r = 0; r < kPer; ++r)
        {
            const int i = lane + 64 * r;
            ox[r] = i < (int)n ? kx[i] : ~0ull;
            oy[r] = i < (int)n ? ky[i] : ~0ull;
            rk[r] = 0;
        }
        for (int j = 0; j < (int)n; ++j)
        {
            const uint64_t jx = kx[j], jy = ky[j];
#pragma unroll
            for (int r = 0; r < kPer; ++r)
            {
                const int i = lane + 64 * r;
                rk[r] += jx < ox[r] || (jx == ox[r] && (jy < oy[r] || (jy == oy[r] && j < i))) ? 1 : 0;
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kPer; ++r)
        {
            const int i = lane + 64 * r;
            if (i < (int)n) ord[rk[r]] = i;
        }
    }
    __syncthreads();
    // runs of equal points: each run's first entry (its lowest index) is the
    // representative; kPer consecutive sorted positions per lane
    int m = 0;
    {
        int head[kPer], cnt = 0;
#pragma unroll
        for (int r = 0; r < kPer; ++r)
        {
            const int i = kPer * lane + r;
            head[r] = 0;
            if (i < (int)n)
            {
                const int v = ord[i];
                head[r] = i == 0 || kx[ord[i - 1]] != kx[v] || ky[ord[i - 1]] != ky[v] ? 1 : 0;
            }
            cnt += head[r];
        }
        const int incl = wave_incl_sum(cnt);
        m = __shfl(incl, 63, 64);
        int d = incl - cnt - 1; // distinct index of the last head before this lane's positions
        // the representative of a run that starts before this lane's
        // positions: the input index at its head, found by a max scan of
        // head positions
        int lasthead = -1;
#pragma unroll
        for (int r = 0; r < kPer; ++r)
            if (head[r]) lasthead = kPer * lane + r;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1)
        {
            const int t = __shfl_up(lasthead, o, 64);
            if (lane >= o) lasthead = max(lasthead, t);
        }
        int runhead = __shfl_up(lasthead, 1, 64);
        if (lane == 0) runhead = -1;
#pragma unroll
        for (int r = 0; r < kPer; ++r)
        {
            const int i = kPer * lane + r;
            if (i >= (int)n) continue;
            const int v = ord[i];
            if (head[r])
            {
                ++d;
                runhead = i;
                ID0[d] = (uint32_t)v;
                P0[d] = pts[v];
                B0[d] = (uint32_t)d;
                dv[d] = v;
            }
            rep[v] = (uint32_t)ord[runhead];
        }
    }
    __syncthreads();
    // frame-1 ranks of the distinct points (y ascending, x descending), by
    // counting as above
    {
        uint64_t ox[kPer], oy[kPer];
        int rk[kPer];
#pragma unroll
        for (int r = 0; r < kPer; ++r)
        {
            const int d = lane + 64 * r;
            const int v = d < m ? dv[d] : 0;
            ox[r] = kx[v];
            oy[r] = ky[v];
            rk[r] = 0;
        }
        for (int j = 0; j < m; ++j)
        {
            const int v = dv[j];
            const uint64_t jx = kx[v], jy = ky[v];
#pragma unroll
            for (int r = 0; r < kPer; ++r) rk[r] += jy < oy[r] || (jy == oy[r] && jx > ox[r]) ? 1 : 0;
        }
#pragma unroll
        for (int r = 0; r < kPer; ++r)
        {
            const int d = lane + 64 * r;
            if (d < m) rank1[d] = (uint32_t)rk[r];
        }
    }
    // a forest of one node at level 0 holding the m distinct points
    if (lane == 0)
    {
        off[0] = 1;
        off[1] = 0;
        off[2] = 0;
        off[kLvlHdr] = 0;
        off[kLvlHdr + 1] = m;
        s_flags = 0;
    }
    __threadfence();
    __syncthreads();
    if (!cell_build<LQE, kPer>(cell_lds, kCap, 0, NX, OR, P0, VE, B0, rank1, off, 0, 0, res, err, counters))
    {
        __syncthreads();
        cell_build<LQX, kPer>(cell_lds, kCap, 0, NX, OR, P0, VE, B0, rank1, off, 0, 0, res, err, counters);
        if (lane == 0) s_flags = 2;
    }
    __threadfence();
    __syncthreads();
    // faces, record by record (64 at a time, so triangles come out in record
    // order): the left faces of edges 4k and 4k+2 that are 3-cycles whose
    // smallest edge is the tested one and that turn counter-clockwise
    // (builder.hxx:98-116).  The cell's records and points are still in LDS
    // (local ids: the distinct points' indices, the cell being all of them).
    const CellLds W(cell_lds, kCap);
    const short* LNX = W.LNX;
    const short* LOR = W.LOR;
    const double2* LP = W.P;
    const int nrec = 3 * m, lim_e = 4 * nrec;
    uint32_t T = 0;
    unsigned nexact = 0;
    for (int k0 = 0; k0 < nrec; k0 += 64)
    {
        const int k = k0 + lane;
        int tri[2][3];
        int c = 0;
        if (k < nrec && LOR[2 * k] >= 0)
            for (int h = 0; h < 2; ++h)
            {
                const int e = 4 * k + 2 * h;
                const int f = e_rot(LNX[e_invrot(e)]);
                if ((unsigned)f >= (unsigned)lim_e || f < e) continue;
                const int g = e_rot(LNX[e_invrot(f)]);
                if ((unsigned)g >= (unsigned)lim_e || g < e) continue;
                if (e_rot(LNX[e_invrot(g)]) != e) continue;
                const int a = LOR[e >> 1], b = LOR[f >> 1], cc = LOR[g >> 1];
                if ((unsigned)a >= (unsigned)m || (unsigned)b >= (unsigned)m || (unsigned)cc >= (unsigned)m) continue;
                const double2 A = LP[a], B = LP[b], C = LP[cc];
                if (!(orient2d(A.x, A.y, B.x, B.y, C.x, C.y, nexact) > 0)) continue;
                tri[c][0] = a;
                tri[c][1] = b;
                tri[c][2] = cc;
                ++c;
            }
        const int incl = wave_incl_sum(c);
        uint32_t pos = T + (uint32_t)(incl - c);
        for (int t = 0; t < c; ++t, ++pos)
            if (pos < tri_cap)
            {
                tris[3 * pos] = ID0[tri[t][0]];
                tris[3 * pos + 1] = ID0[tri[t][1]];
                tris[3 * pos + 2] = ID0[tri[t][2]];
            }
        T += (uint32_t)__shfl(incl, 63, 64);
    }
    if (lane == 0)
    {
        out[0] = T;
        out[1] = (uint32_t)m;
        out[2] = (uint32_t)s_flags;
        if (T > tri_cap) atomicOr(err, 1 << 22);
    }
}

// ---------------------------------------------------------------------------
// 7b: large merges cut into seam pieces (see seam.hpp)
// ---------------------------------------------------------------------------

constexpr int kSegMaxK = 512;          // anchor slots per node (a power of two)
constexpr int kSegCap = kSegMaxK - 2;  // anchors per node at most: a phase's pieces fit 256 threads
constexpr int kSegSlots = kSegMaxK + 1; // pieces per node at most
constexpr int kSegThreads = 256;       // threads of the per-node compaction and joining kernels
constexpr int kSegMin = 1024;          // smaller nodes merge in one piece
// A level merges one node per lane (k_merge) instead of in seam pieces when
// its longest seam asks for fewer than kSegLevelAnchors anchors (the pieces'
// per-level kernels cost more than the chains they shorten), or when it has
// a merge for every lane of k_merge's layout on the GPU (256 CUs x 4 SIMDs x
// 16 lanes) and no node of kSegLevelBig points or more.
constexpr int kSegLevelAnchors = 8;
constexpr int kSegLevelNodes = 16384;
constexpr int kSegLevelBig = 4096;
constexpr int kPhases = 2;             // pieces k, k+2, ... run together
constexpr int kSeamErr = 64;           // err bit: redo the run with unsplit merges

// Anchors for a node of s points whose extent across the cut is w and along
// it h: the seam has about sqrt(s h / w) cross edges (the points within one
// spacing of the cut line), and a piece gets ~3.5 of them.  0: one piece.
__host__ __device__ inline int seam_anchors(double s, double w, double h)
{
    if (s < kSegMin) return 0;
    const double seam = w > 0 ? std::sqrt(s * h / w) : 1e9;
    const double k = seam / 3.5;
    if (k < 3) return 0;
    return k > kSegCap ? kSegCap : (int)k;
}

// upper bound of seam_anchors for nodes of up to s points (elongation up to
// 16:1 along the cut)
__host__ __device__ inline int seam_anchors_max(double s)
{
    return s < kSegMin ? 0 : seam_anchors(s, 1.0, 16.0);
}

// split vertex of every internal node; anchor heights of segmented levels
__global__ __launch_bounds__(kBlock) void k_node_info(const uint32_t* Bf, const uint32_t* Bo, const uint32_t* Kf,
                                                      const int* off, int d, int K, const double2* PT, int* splitv,
                                                      int* qv, int* kn, int* lvlK, uint32_t* medk, int2* ext)
{
    const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= lcount(off, d)) return;
    const int* od = off + off_base(off, d);
    const int o = od[j];
    const int s = od[j + 1] - o;
    medk[j] = s < 4 ? 0xffffffffu : Kf[o + left_size(s)]; // a node of < 4 points passes through
    if (s < 4) return;
    const size_t ni = node_index(off, d, j);
    splitv[ni] = (int)Bf[o + left_size(s)];
    // the node's first and last points in the other frame, which is its
    // parent's: the ends its merged hull edges are re-walked to
    ext[ni] = make_int2((int)Bo[o], (int)Bo[o + s - 1]);
    int Kn = 0;
    if (K)
    {
        // extents from the two orders' ends: across the cut (this level's
        // frame) and along it (the other frame)
        const int axis = d & 1;
        const double2 f0 = PT[Bf[o]], f1 = PT[Bf[o + s - 1]];
        const double2 g0 = PT[Bo[o]], g1 = PT[Bo[o + s - 1]];
        const double w = axis ? f1.y - f0.y : f1.x - f0.x;
        const double h = axis ? g1.x - g0.x : g1.y - g0.y;
        Kn = min(K, seam_anchors((double)s, w, h));
    }
    kn[ni] = Kn;
    if (Kn) atomicMax(&lvlK[d], Kn);
}

// anchor heights: the vertex at the middle of the k-th of Kn equal slices of
// the node's order along the cut (one thread per (node, k))
__global__ __launch_bounds__(kBlock) void k_node_qv(const uint32_t* Bo, const int* off, int d, int K, const int* kn,
                                                    int* qv)
{
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= lcount(off, d) * (uint32_t)K) return;
    const uint32_t j = t / K;
    const int k = (int)(t % K);
    const int* od = off + off_base(off, d);
    const int o = od[j];
    const int s = od[j + 1] - o;
    if (s < 4) return;
    const size_t ni = node_index(off, d, j);
    const int Kn = kn[ni];
    if (k < Kn) qv[ni * kSegMaxK + k] = (int)Bo[o + (int)(((long long)(2 * k + 1) * s) / (2 * Kn))];
}

__device__ __forceinline__ bool less_frame_pt(double2 a, double2 b, int frame)
{
    if (frame == 0) return a.x < b.x || (a.x == b.x && a.y < b.y);
    return a.y < b.y || (a.y == b.y && a.x > b.x);
}

// A vertex of the deepest built kd node whose region contains p, below node
// (dd, jj): the node's point closest to the cut of frame `axis` on the given
// side (0: the lower child, which faces the cut with its largest points in
// that frame; 1: the upper child), so that the walks toward p start close.
__device__ int kd_descend(const int* off, const int* splitv, const double2* PT, const uint32_t* B0,
                          const uint32_t* B1, const uint32_t* newid, int D, int dd, uint32_t jj, double2 p, int axis,
                          int side)
{
    while (dd < D)
    {
        const int* od = off + off_base(off, dd);
        const int s = od[jj + 1] - od[jj];
        if (s <= 3) break;
        const double2 sp = PT[splitv[node_index(off, dd, jj)]];
        jj = 2 * jj + (less_frame_pt(p, sp, dd & 1) ? 0u : 1u);
        ++dd;
    }
    const int* od = off + off_base(off, dd);
    const int o = od[jj], e = od[jj + 1];
    if (e <= o) return -1;
    const uint32_t* B = axis ? B1 : B0;
    const uint32_t v = B[side ? o : e - 1];
    return (int)(axis && newid ? newid[v] : v); // B1 holds ids of the order before renumbering
}

constexpr int kCross = 16;                     // crossing edges per side of an anchor at most
// Work bounds of one anchor.  An anchor is optional (a piece simply runs
// longer without it) while a level's anchor kernel lasts as long as its
// slowest anchor, so the walks toward the query point are cut short: steps
// of the first walk from the kd descent's vertex, steps of each walk after
// the point moves to the bisector, and bisector moves.  Over 99.7% of the
// anchors finish within these (10M uniform and Gaussian inputs); the rest,
// whose walks run for hundreds of steps, were the kernel's tail.
constexpr int kAnchorWalk = 24;
constexpr int kAnchorRewalk = 8;
constexpr int kAnchorMoves = 3;
constexpr int kRec = 6 + 2 * kCross;           // ints per anchor record: u v xu xv nl nr crossings
constexpr int kMarkSlotsMax = 4096;            // anchors per level up to which crossings get a thread each

// Two lanes per anchor (adjacent, so in one wavefront): lane side 0 works in
// the lower child L, lane side 1 in the upper child R.  Every decision both
// lanes take depends only on values they exchange, so a pair never diverges
// and its shuffles always find the partner active; the two walks, ring scans
// and dry runs of an anchor overlap their memory latencies.  One more thread
// per node, after all the anchors' lanes, re-walks the children's hull
// edges to this level's frame (child_result) and walks the node's lower
// common tangent (reads only, like the dry runs; its result goes to nstate,
// and k_seg_select makes the connect), so those walks overlap the anchors.
__global__ __launch_bounds__(kMergeBlock) void k_seg_anchor(int* NX, int* OR, const double2* PT, int* VE,
                                                            const int* off, const int* splitv, const int* qv,
                                                            const uint32_t* B0, const uint32_t* B1,
                                                            const uint32_t* newid, const int* kn,
                                                            int D, int d, int K, uint32_t m,
                                                            int lim_e, int* arec, double4* acoord, double* akey,
                                                            int* astate, int4* res, const int2* ext, int4* nstate,
                                                            int* err, unsigned long long* counters)
{
    const uint32_t lane2 = blockIdx.x * kMergeBlock + threadIdx.x;
    const uint32_t t = lane2 >> 1; // anchor
    const int side = (int)(lane2 & 1);
    unsigned long long certified = 0, dry_ok = 0;
    GQE q;
    q.init(NX, OR, PT, VE, lim_e, (int)m);
    const uint32_t nanch = lcount(off, d) * (uint32_t)K;
    if (lane2 >= 2 * nanch && lane2 < 2 * nanch + lcount(off, d))
    {
        const uint32_t j = lane2 - 2 * nanch;
        const int* od = off + off_base(off, d);
        const int s = od[j + 1] - od[j];
        if (s >= 4)
        {
            const int4 L = child_result(q, res, ext, off, d + 1, 2 * j);
            const int4 R = child_result(q, res, ext, off, d + 1, 2 * j + 1);
            int ldi = L.y, rdi = R.x;
            int budget = 16 * s + 64;
            q.tangent_walk(ldi, rdi, budget);
            nstate[j] = make_int4(ldi, rdi, 0, 0);
            if (q.err) atomicOr(err, 2 << q.err);
        }
    }
    else if (t < nanch)
    {
        const uint32_t j = t / K;
        const int k = (int)(t % K);
        int state = 0;
        const int* od = off + off_base(off, d);
        const int s = od[j + 1] - od[j];
        if (s >= 4 && k < kn[node_index(off, d, j)])
        {
            const int axis = d & 1;
            const size_t ni = node_index(off, d, j);
            const double c = frame_x(PT[splitv[ni]], axis);
            const double Y = frame_y(PT[qv[ni * kSegMaxK + k]], axis);
            double2 p = from_frame(c, Y, axis);
            // own side: x (nearest vertex) and ex (an edge out of it)
            const int st = kd_descend(off, splitv, PT, B0, B1, newid, D, d + 1, 2 * j + side, p, axis, side);
            int ex = -1;
            int x = st >= 0 ? nn_walk(q, st, p, kAnchorWalk, ex) : -1;
            int xo = __shfl_xor(x, 1), exo = __shfl_xor(ex, 1);
            int u = side ? xo : x, v = side ? x : xo;
            for (int it = 0; it < kAnchorMoves && u >= 0 && v >= 0; ++it)
            {
                if (q.err | __shfl_xor(q.err, 1)) break;
                const double2 U = PT[u], V = PT[v];
                const double ux = frame_x(U, axis), uy = frame_y(U, axis);
                const double vx = frame_x(V, axis), vy = frame_y(V, axis);
                if (!(vx > ux)) break;
                const double X = ((vx - ux) * (vx + ux) + (Y - vy) * (Y - vy) - (Y - uy) * (Y - uy)) / (2.0 * (vx - ux));
                p = from_frame(X, Y, axis);
                int ex2 = -1;
                const int x2 = nn_walk(q, x, p, kAnchorRewalk, ex2);
                const int xo2 = __shfl_xor(x2, 1), exo2 = __shfl_xor(ex2, 1);
                const bool same = x2 == x && xo2 == xo;
                x = x2;
                xo = xo2;
                ex = ex2;
                exo = exo2;
                u = side ? xo : x;
                v = side ? x : xo;
                if (same) break;
            }
            const int eu = side ? exo : ex, ev = side ? ex : exo;
            bool ok = u >= 0 && v >= 0 && eu >= 0 && ev >= 0 && !(q.err | __shfl_xor(q.err, 1));
            if (ok)
            {
                int wl = -1, wr = -1;
                const int ring = certify_ring(q, u, v, side ? ev : eu, wl, wr) ? 1 : 0;
                const int wlo = __shfl_xor(wl, 1), wro = __shfl_xor(wr, 1), ringo = __shfl_xor(ring, 1);
                ok = ring && ringo && !(q.err | __shfl_xor(q.err, 1));
                if (ok)
                    ok = certify_join(q, u, v, side ? wlo : wl, side ? wro : wr, side ? wl : wlo, side ? wr : wro);
            }
            if (ok)
            {
                if (side == 0) certified = 1;
                // dry run of the insertion on the unmerged halves: each lane
                // finds its end's wedge and the crossings on its side
                int* r = arec + (size_t)t * kRec;
                const int me = side ? v : u, other = side ? u : v;
                const int xw = find_wedge(q, me, side ? ev : eu, other);
                const int n = xw >= 0 ? collect_crossings(q, me, xw, other, r + 6 + side * kCross, kCross) : -1;
                const int no = __shfl_xor(n, 1);
                const int bad = q.err | __shfl_xor(q.err, 1);
                if (n >= 0 && no >= 0 && !bad)
                {
                    r[side] = me;
                    r[2 + side] = xw;
                    r[4 + side] = n;
                    if (side == 0)
                    {
                        const double2 U = PT[u], V = PT[v];
                        acoord[t] = make_double4(U.x, U.y, V.x, V.y);
                        // height where the anchor crosses the cut line: the sort key
                        const double ux = frame_x(U, axis), uy = frame_y(U, axis);
                        const double vx = frame_x(V, axis), vy = frame_y(V, axis);
                        akey[t] = vx > ux ? uy + (c - ux) / (vx - ux) * (vy - uy) : 0.5 * (uy + vy);
                        dry_ok = 1;
                    }
                    state = 1;
                }
            }
        }
        if (side == 0) astate[t] = state;
    }
    wave_add(&counters[5], certified);
    wave_add(&counters[10], dry_ok);
    flush_counts(counters, q.pc);
}

// Anchors in chain order, each strictly above the previous cut, and no two
// cuts with a common end vertex.  One workgroup per node (a power of two of
// threads, at least the node's candidate count): the candidates and their
// cached coordinates go to LDS and are sorted by crossing height (bitonic).
// The choice is the greedy one: in height order a candidate is taken when
// neither end is an end of the tangent or of a cut taken before it, and both
// ends lie strictly left of the last cut taken (exact orient2d).  The tangent's
// rings changed after the anchors' dry runs, so its two vertices are excluded
// as anchor ends too.
//
// The greedy pass runs in parallel as a fixed-point iteration.  Each round,
// every position recomputes its decision from the previous round's decisions
// of the positions before it: the last taken position comes from a prefix
// max, and "an end already used" from an LDS table holding, per vertex, the
// lowest taken position that has it as an end.  A position's decision depends
// only on earlier positions, so after round r the first r positions are
// final; the rounds stop when nothing changes (two or three rounds when
// rejections are rare, as they are).
constexpr int kSelHash = 2048; // > 2 * (kSegMaxK + 1), a power of two

// per-vertex lowest taken position: insert (v, pos) / look up (INT_MAX: none)
__device__ __forceinline__ void vmin_put(int* hk, int* hp, int v, int pos)
{
    unsigned x = ((unsigned)v * 2654435761u) & (kSelHash - 1);
    for (int probe = 0; probe < kSelHash; ++probe)
    {
        int cur = hk[x];
        if (cur < 0)
        {
            cur = atomicCAS(&hk[x], -1, v);
            if (cur < 0) cur = v;
        }
        if (cur == v)
        {
            atomicMin(&hp[x], pos);
            return;
        }
        x = (x + 1) & (kSelHash - 1);
    }
}

__device__ __forceinline__ int vmin_get(const int* hk, const int* hp, int v)
{
    unsigned x = ((unsigned)v * 2654435761u) & (kSelHash - 1);
    for (int probe = 0; probe < kSelHash; ++probe)
    {
        const int cur = hk[x];
        if (cur == v) return hp[x];
        if (cur < 0) return INT_MAX;
        x = (x + 1) & (kSelHash - 1);
    }
    return INT_MAX;
}

// inclusive block scan (max or sum) for blocks of a multiple of 64 threads
template <bool kMax>
__device__ __forceinline__ int sel_scan(int v, int* wtmp)
{
    const int lane = (int)__lane_id(), w = (int)threadIdx.x >> 6;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1)
    {
        const int t = __shfl_up(v, o, 64);
        if (lane >= o) v = kMax ? max(v, t) : v + t;
    }
    if (lane == 63) wtmp[w] = v;
    __syncthreads();
    int pre = kMax ? -1 : 0;
    for (int i = 0; i < w; ++i) pre = kMax ? max(pre, wtmp[i]) : pre + wtmp[i];
    __syncthreads();
    return kMax ? max(pre, v) : pre + v;
}

__global__ __launch_bounds__(kSegMaxK) void k_seg_select(int* NX, int* OR, const double2* PT, int* VE, const int* off,
                                                         int d, int K, uint32_t m, int lim_e, unsigned* spare,
                                                         int spare_base, int spare_cap, const int4* res,
                                                         int4* nstate, int* segstart, const int* arec,
                                                         const double4* acoord, const double* akey, int* astate,
                                                         int* selk, int* nsel, int* alist_n, int* err,
                                                         unsigned long long* counters)
{
    __shared__ double skey[kSegMaxK];
    __shared__ int sslot[kSegMaxK];
    __shared__ int su[kSegMaxK], sv[kSegMaxK];
    __shared__ double4 sc[kSegMaxK];
    __shared__ int hk[kSelHash], hp[kSelHash];
    __shared__ int lastv[kSegMaxK];
    __shared__ int wtmp[kSegMaxK / 64];
    const uint32_t j = blockIdx.x;
    const int k = threadIdx.x, nt = blockDim.x;
    if (j == 0 && k == 0) *alist_n = 0;
    if (j >= lcount(off, d)) return;
    const int* od = off + off_base(off, d);
    const int s = od[j + 1] - od[j];
    if (k == 0) nsel[j] = 0;
    if (s < 4) return;
    if (k == 0)
    {
        // the lower common tangent (walked by the anchor kernel) connected:
        // the first cross edge, and the parent's hull edges and pool
        GQE q;
        q.init(NX, OR, PT, VE, lim_e, (int)m);
        q.spare = spare;
        q.spare_base = spare_base;
        q.spare_cap = spare_cap;
        const int4 L = res[node_index(off, d + 1, 2 * j)];
        const int4 R = res[node_index(off, d + 1, 2 * j + 1)];
        Pool pool = q.cat(Pool{ L.z, L.w }, Pool{ R.z, R.w });
        int ldo = L.x, rdo = R.y;
        const int4 w = nstate[j];
        segstart[(size_t)j * kSegSlots] = q.tangent_connect(ldo, w.x, w.y, rdo, pool);
        nstate[j] = make_int4(ldo, rdo, pool.head, pool.tail);
        flush_counts(counters, q.pc);
        if (q.err) atomicOr(err, 2 << q.err);
    }
    __syncthreads();
    const int b0 = segstart[(size_t)j * kSegSlots];
    if (b0 < 0 || b0 >= lim_e) return; // the tangent failed: the run's error flag is set
    const int pu0 = OR[(b0 ^ 2) >> 1], pv0 = OR[b0 >> 1];
    if ((unsigned)pu0 >= m || (unsigned)pv0 >= m) return;
    {
        const size_t t = (size_t)j * K + k;
        const bool valid = k < K && astate[t] == 1;
        skey[k] = valid ? akey[t] : INFINITY;
        sslot[k] = valid ? k : -1;
        if (valid)
        {
            su[k] = arec[t * kRec];
            sv[k] = arec[t * kRec + 1];
            sc[k] = acoord[t];
        }
    }
    __syncthreads();
    // bitonic sort of (key, slot), ascending; invalid slots (infinite keys) last
    for (int size = 2; size <= nt; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1)
        {
            const int partner = k ^ stride;
            if (partner > k)
            {
                const bool up = (k & size) == 0;
                const double a = skey[k], b = skey[partner];
                const int sa = sslot[k], sb = sslot[partner];
                // ties (and invalid pairs) by slot, so the order is a total one
                const bool gt = a > b || (a == b && sa > sb);
                if (gt == up)
                {
                    skey[k] = b;
                    skey[partner] = a;
                    sslot[k] = sb;
                    sslot[partner] = sa;
                }
            }
            __syncthreads();
        }
    // the candidate at sorted position k, its tests against the tangent and
    // against the candidate at position k - 1
    const int sl = sslot[k];
    int u = -1, v = -1;
    double4 c = make_double4(0, 0, 0, 0);
    bool base = false, ok_tan = false, ok_prev = false;
    unsigned dummy = 0;
    if (sl >= 0)
    {
        u = su[sl];
        v = sv[sl];
        c = sc[sl];
        base = u != v && u != pu0 && u != pv0 && v != pu0 && v != pv0;
        const double2 PU = PT[pu0], PV = PT[pv0];
        ok_tan = orient2d(c.x, c.y, PU.x, PU.y, PV.x, PV.y, dummy) > 0 &&
            orient2d(c.z, c.w, PU.x, PU.y, PV.x, PV.y, dummy) > 0;
        const int sp = k > 0 ? sslot[k - 1] : -1;
        if (sp >= 0)
        {
            const double4 q = sc[sp];
            ok_prev = orient2d(c.x, c.y, q.x, q.y, q.z, q.w, dummy) > 0 &&
                orient2d(c.z, c.w, q.x, q.y, q.z, q.w, dummy) > 0;
        }
    }
    bool take = base;
    for (int round = 0; round <= nt; ++round)
    {
        lastv[k] = sel_scan<true>(take ? k : -1, wtmp);
        for (int x = k; x < kSelHash; x += nt)
        {
            hk[x] = -1;
            hp[x] = INT_MAX;
        }
        __syncthreads();
        if (take)
        {
            vmin_put(hk, hp, u, k);
            vmin_put(hk, hp, v, k);
        }
        __syncthreads();
        bool now = false;
        if (base && vmin_get(hk, hp, u) >= k && vmin_get(hk, hp, v) >= k)
        {
            const int L = k > 0 ? lastv[k - 1] : -1;
            if (L < 0)
                now = ok_tan;
            else if (L == k - 1)
                now = ok_prev;
            else
            {
                const double4 q = sc[sslot[L]];
                now = orient2d(c.x, c.y, q.x, q.y, q.z, q.w, dummy) > 0 &&
                    orient2d(c.z, c.w, q.x, q.y, q.z, q.w, dummy) > 0;
            }
        }
        const bool changed = __syncthreads_or(now != take);
        take = now;
        if (!changed) break;
    }
    const int rank = sel_scan<false>(take ? 1 : 0, wtmp);
    if (sl >= 0)
    {
        const size_t tt = (size_t)j * K + sl;
        astate[tt] = take ? 2 : 0;
        if (take) selk[(size_t)j * kSegSlots + rank - 1] = (int)tt;
    }
    if (k == nt - 1)
    {
        nsel[j] = rank;
        if (rank) atomicAdd(&counters[11], (unsigned long long)rank);
    }
}

// Parallel anchor insertion.  Several anchors may cross the same edge (long
// edges along a hull are crossed by many), so insertion is done per vertex
// rather than per anchor: every edge crossed by a selected anchor is marked
// for deletion (del[record] = the anchor that frees it), every vertex whose
// ring changes is listed once, and one thread per listed vertex rewrites its
// ring: the old cyclic order with the marked edges dropped and the new anchor
// edge placed right after its wedge edge.  A ring fixes the onext of its own
// directed edges and of their rotations (rot(e).onext = invrot(oprev(e))),
// so rings of distinct vertices touch disjoint words of NX.

// anchor edge records, vertex marks for u and v (owner[w] = anchor index)
__global__ __launch_bounds__(kMergeBlock) void k_seg_mark(int* OR, int total, const int* arec, const int* astate,
                                                          unsigned* spare, int spare_base, int spare_cap, int* owner,
                                                          int* aedge, int* alist, int* alist_n, int acap, int* err,
                                                          unsigned long long* counters)
{
    const int t = blockIdx.x * kMergeBlock + threadIdx.x;
    const bool act = t < total && astate[t] == 2;
    const int* r = arec + (size_t)(act ? t : 0) * kRec;
    // one record per inserted anchor, one allocation per wavefront
    const unsigned k = wave_fetch_add(spare, act ? 1u : 0u);
    int e = -1;
    if (act)
    {
        if ((int)k >= spare_cap)
        {
            atomicOr(err, kSeamErr);
            aedge[t] = -1;
        }
        else
        {
            e = 4 * (spare_base + (int)k); // e runs v -> u
            OR[e >> 1] = r[1];
            OR[(e ^ 2) >> 1] = r[0];
            aedge[t] = e;
            owner[r[0]] = t;
            owner[r[1]] = t;
        }
    }
    const int at = wave_fetch_add(alist_n, e >= 0 ? 2 : 0);
    if (e >= 0)
    {
        if (at + 2 <= acap)
        {
            alist[at] = r[0];
            alist[at + 1] = r[1];
        }
        else
            atomicOr(err, kSeamErr);
    }
    wave_add(&counters[6], e >= 0 ? 1ull : 0ull);
}

// crossed edges: deletion marks and the rings of their endpoints.  A
// thread's crossing slots are independent, so they go in chunks of kMarkChunk
// per side: the chunk's edge ids, then their end vertices, then the
// deletion marks and the vertex claims are issued before any result is
// waited on (a thread waits on a few round trips to L2 instead of four per
// slot).  Every lane runs as many chunks as the longest list in its
// wavefront, and the vertex list grows by one atomic per wavefront and chunk.
constexpr int kMarkChunk = 4;

__global__ __launch_bounds__(kMergeBlock) void k_seg_mark_cross(const int* OR, int total, const int* arec,
                                                                const int* astate, const int* aedge, uint32_t m,
                                                                int lim_e, int* owner, int* del, int* alist,
                                                                int* alist_n, int acap, int* err)
{
    const int t = blockIdx.x * kMergeBlock + threadIdx.x;
    const bool act = t < total && astate[t] == 2 && aedge[t] >= 0;
    const int* r = arec + (size_t)(act ? t : 0) * kRec;
    const int c0 = act ? min(r[4], kCross) : 0, c1 = act ? min(r[5], kCross) : 0;
    int cmax = max(c0, c1);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) cmax = max(cmax, __shfl_xor(cmax, o, 64));
    for (int i0 = 0; i0 < cmax; i0 += kMarkChunk)
    {
        int e[2 * kMarkChunk], w[4 * kMarkChunk];
#pragma unroll
        for (int x = 0; x < 2 * kMarkChunk; ++x)
        {
            const int side = x / kMarkChunk, i = i0 + x % kMarkChunk;
            e[x] = i < (side ? c1 : c0) ? r[6 + side * kCross + i] : -1;
        }
#pragma unroll
        for (int x = 0; x < 2 * kMarkChunk; ++x)
        {
            const bool ok = e[x] >= 0 && e[x] < lim_e;
            w[2 * x] = ok ? OR[e[x] >> 1] : -1;
            w[2 * x + 1] = ok ? OR[(e[x] ^ 2) >> 1] : -1;
        }
        bool bad = false;
#pragma unroll
        for (int x = 0; x < 2 * kMarkChunk; ++x)
        {
            const int side = x / kMarkChunk, i = i0 + x % kMarkChunk;
            if (i < (side ? c1 : c0) && !(e[x] >= 0 && e[x] < lim_e)) bad = true;
            if (e[x] >= 0 && e[x] < lim_e) atomicCAS(&del[e[x] >> 2], -1, t);
        }
        if (bad) atomicOr(err, kSeamErr);
        int got[4 * kMarkChunk];
#pragma unroll
        for (int y = 0; y < 4 * kMarkChunk; ++y)
            got[y] = (unsigned)w[y] < m ? atomicCAS(&owner[w[y]], -1, -3) : 0;
        int want = 0;
#pragma unroll
        for (int y = 0; y < 4 * kMarkChunk; ++y)
        {
            if ((unsigned)w[y] >= m || got[y] != -1) w[y] = -1;
            want += w[y] >= 0 ? 1 : 0;
        }
        int at = wave_fetch_add(alist_n, want);
#pragma unroll
        for (int y = 0; y < 4 * kMarkChunk; ++y)
        {
            if (w[y] < 0) continue;
            if (at < acap)
                alist[at] = w[y];
            else
                atomicOr(err, kSeamErr);
            ++at;
        }
    }
}

// The same with one thread per (anchor, side, crossing slot), for levels of
// few anchors: an anchor's atomics then do not wait on each other (the vertex
// list grows by one atomic per wavefront).
__global__ __launch_bounds__(kMergeBlock) void k_seg_mark_cross_slots(const int* OR, int total, const int* arec,
                                                                const int* astate, const int* aedge, uint32_t m,
                                                                int lim_e, int* owner, int* del, int* alist,
                                                                int* alist_n, int acap, int* err)
{
    const size_t x = (size_t)blockIdx.x * kMergeBlock + threadIdx.x;
    const size_t t = x / (2 * kCross);
    const int side = (int)(x / kCross) & 1, i = (int)(x % kCross);
    bool has = false;
    int e = 0;
    if (t < (size_t)total && astate[t] == 2 && aedge[t] >= 0)
    {
        const int* r = arec + t * kRec;
        if (i < r[4 + side])
        {
            has = true;
            e = r[6 + side * kCross + i];
        }
    }
    int w0 = -1, w1 = -1;
    if (has && (e < 0 || e >= lim_e))
        atomicOr(err, kSeamErr);
    else if (has)
    {
        atomicCAS(&del[e >> 2], -1, (int)t);
        const int a = OR[e >> 1], b = OR[(e ^ 2) >> 1];
        if ((unsigned)a < m && atomicCAS(&owner[a], -1, -3) == -1) w0 = a;
        if ((unsigned)b < m && atomicCAS(&owner[b], -1, -3) == -1) w1 = b;
    }
    // the whole wavefront takes its list slots together
    const int want = (w0 >= 0 ? 1 : 0) + (w1 >= 0 ? 1 : 0);
    int at = wave_fetch_add(alist_n, want);
    for (int w : { w0, w1 })
    {
        if (w < 0) continue;
        if (at < acap)
            alist[at] = w;
        else
            atomicOr(err, kSeamErr);
        ++at;
    }
}

// one thread per listed vertex: its new ring
__global__ __launch_bounds__(kMergeBlock) void k_seg_rebuild(int* NX, const int* OR, int* VE, uint32_t m, int lim_e,
                                                             const int* arec, const int* aedge, int* owner,
                                                             const int* del, const int* alist, const int* alist_n,
                                                             int acap, int* err, unsigned long long* counters)
{
    const int n = min(*alist_n, acap);
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&counters[12], (unsigned long long)n);
    for (int i = blockIdx.x * kMergeBlock + threadIdx.x; i < n; i += gridDim.x * kMergeBlock)
    {
        const int w = alist[i];
        if ((unsigned)w >= m) continue;
        const int t = owner[w];
        int ins = -1, after = -1;
        if (t >= 0)
        {
            const int* r = arec + (size_t)t * kRec;
            const int e = aedge[t];
            if (w == r[0])
            {
                ins = e ^ 2; // u -> v
                after = r[2];
            }
            else
            {
                ins = e; // v -> u
                after = r[3];
            }
        }
        const int e0 = VE[w];
        if (e0 < 0 || e0 >= lim_e)
        {
            atomicOr(err, kSeamErr);
            continue;
        }
        int first = -1, prev = -1, cur = e0, guard = 0;
        bool placed = ins < 0;
        do
        {
            const int nxt = NX[cur]; // read before any write to NX[cur]
            if (del[cur >> 2] == -1)
            {
                if (prev >= 0)
                {
                    NX[prev] = cur;
                    NX[e_rot(cur)] = e_invrot(prev);
                }
                else
                    first = cur;
                prev = cur;
            }
            if (cur == after)
            {
                if (prev >= 0)
                {
                    NX[prev] = ins;
                    NX[e_rot(ins)] = e_invrot(prev);
                }
                else
                    first = ins;
                prev = ins;
                placed = true;
            }
            cur = nxt;
            if (cur < 0 || cur >= lim_e) break;
        } while (cur != e0 && ++guard < 65536);
        if (cur != e0 || !placed || first < 0)
        {
            atomicOr(err, kSeamErr); // malformed ring, wedge edge gone, or an isolated vertex
            owner[w] = -1;
            continue;
        }
        NX[prev] = first;
        NX[e_rot(first)] = e_invrot(prev);
        VE[w] = first;
        owner[w] = -1;
    }
}

// freed records onto the private list of the anchor that marked them first
__global__ __launch_bounds__(kMergeBlock) void k_seg_pools(int* NX, int* OR, int total, const int* arec,
                                                           const int* astate, const int* aedge, int lim_e, int* del,
                                                           int2* apool)
{
    const int t = blockIdx.x * kMergeBlock + threadIdx.x;
    if (t >= total || astate[t] != 2) return;
    int head = -1, tail = -1;
    if (aedge[t] >= 0)
    {
        const int* r = arec + (size_t)t * kRec;
        for (int side = 0; side < 2; ++side)
        {
            const int c = r[4 + side];
            const int* ce = r + 6 + side * kCross;
            for (int i = 0; i < c; ++i)
            {
                const int e = ce[i];
                if (e < 0 || e >= lim_e) continue;
                const int k = e >> 2;
                if (del[k] != t) continue;
                del[k] = -1;
                OR[2 * k] = -1;
                NX[4 * k] = head;
                if (head < 0) tail = k;
                head = k;
            }
        }
    }
    apool[t] = make_int2(head, tail);
}

// cut list of every node: the tangent, then the inserted anchors in order
// (one workgroup per node, a block scan compacts the inserted ones)
__global__ __launch_bounds__(kSegThreads) void k_seg_compact(int d, const int* selk, const int* nsel,
                                                            const int* astate, const int* aedge, int* segstart,
                                                            int* nseg, const int* off)
{
    constexpr int kPerSel = kSegMaxK / kSegThreads; // candidates per thread, consecutive
    const uint32_t j = blockIdx.x;
    const int k = threadIdx.x;
    if (j >= lcount(off, d)) return;
    const int* od = off + off_base(off, d);
    const int s = od[j + 1] - od[j];
    if (s < 4)
    {
        if (k == 0) nseg[j] = 0;
        return;
    }
    const int ns = nsel[j];
    int v[kPerSel];
    uint32_t c = 0;
#pragma unroll
    for (int r = 0; r < kPerSel; ++r)
    {
        const int x = k * kPerSel + r;
        v[r] = -1;
        if (x < ns)
        {
            const int t = selk[(size_t)j * kSegSlots + x];
            if (astate[t] == 2 && aedge[t] >= 0) v[r] = aedge[t];
        }
        c += v[r] >= 0 ? 1u : 0u;
    }
    uint32_t total;
    uint32_t pos = block_exclusive_scan(c, total);
#pragma unroll
    for (int r = 0; r < kPerSel; ++r)
        if (v[r] >= 0) segstart[(size_t)j * kSegSlots + 1 + pos++] = v[r];
    if (k == 0) nseg[j] = 1 + (int)total;
}

// The pieces of a level: the rising bubble from cut i to cut i+1, in two
// phases (even pieces, then odd: pieces of one phase share no cut vertex).
// Every piece of a node runs in the same workgroup (tpn threads per node,
// npb nodes per workgroup), so the phases are separated by a workgroup
// barrier and the vertex claims live in an LDS hash table (cleared between
// phases): a vertex belongs to one node of the level, and only that node's
// pieces claim it.
// A piece that needs a vertex another piece of its phase holds stops and
// leaves its resume edge in segresume; after both phases one thread per node
// finishes the stopped pieces bottom to top, with nothing else running on
// the node (the bubble is a pure function of the store, so a resumed piece
// does exactly what an uninterrupted one would).
// Only every `spread`-th lane runs pieces: on levels whose pieces would
// leave SIMDs idle they are spread over more wavefronts with fewer each
// (measured on 10M uniform points: up to 1,024 wavefronts 22.47 -> 22.40 ms
// per run, up to 2,048 or 4,096 slower; later, up to 512 against 1,024:
// 21.55 -> 21.45 ms, 256 the same as 512, Gaussian points unchanged).
constexpr int kSegPieceBlock = 256; // threads per workgroup at most
constexpr int kPieceWaves = 512;    // wavefronts of a spread launch at most

__global__ __launch_bounds__(kSegPieceBlock) void k_seg_pieces(int* NX, int* OR, const double2* PT, int* VE,
                                                              const int* off, int d, uint32_t m, int lim_e,
                                                              unsigned* spare, int spare_base, int spare_cap,
                                                              const int* segstart, const int* nseg, int tpn,
                                                              int npb, int spread, int claim_entries,
                                                              int2* segpool, int* segresume, int* err,
                                                              unsigned long long* counters)
{
    __shared__ int nstop[kSegPieceBlock];
    extern __shared__ unsigned long long claim_table[];
    const int slot = threadIdx.x / spread;
    const int ln = slot / tpn, h = slot % tpn;
    const uint32_t j = blockIdx.x * (uint32_t)npb + (uint32_t)ln;
    const bool node = threadIdx.x % spread == 0 && ln < npb && j < lcount(off, d);
    int ns = 0, s = 0;
    const int* st = nullptr;
    if (node)
    {
        const int* od = off + off_base(off, d);
        s = od[j + 1] - od[j];
        ns = nseg[j];
        st = segstart + (size_t)j * kSegSlots;
    }
    if (h == 0) nstop[ln] = 0;
    unsigned long long ran = 0, stopped = 0, resumed = 0;
    GQE q;
    q.init(NX, OR, PT, VE, lim_e, (int)m);
    q.spare = spare;
    q.spare_base = spare_base;
    q.spare_cap = spare_cap;
    for (int parity = 0; parity < kPhases; ++parity)
    {
        for (int x = threadIdx.x; x < claim_entries; x += blockDim.x) claim_table[x] = kNoClaim;
        __syncthreads();
        for (int i = kPhases * h + parity; node && i < ns; i += kPhases * tpn)
        {
            q.claim = claim_table;
            q.claim_mask = claim_entries - 1;
            q.claim_me = i;
            q.h0 = q.h1 = q.h2 = q.h3 = -1;
            const size_t t = (size_t)j * kSegSlots + i;
            int basel = st[i];
            const int stop = i + 1 < ns ? st[i + 1] : -1;
            Pool pool = { -1, -1 };
            int budget = 16 * s + 64;
            const int r = q.bubble(basel, stop, pool, budget);
            ++ran;
            if (r == 2 && !q.err)
            {
                segresume[t] = basel;
                atomicAdd(&nstop[ln], 1);
                ++stopped;
            }
            else
            {
                segresume[t] = -1;
                if (stop >= 0 && r != 1 && !q.err) q.err = 4;
            }
            segpool[t] = make_int2(pool.head, pool.tail);
        }
        __syncthreads();
    }
    if (node && h == 0 && nstop[ln] > 0)
    {
        // stopped pieces, bottom to top, with claims off
        q.claim = nullptr;
        int budget = 16 * s + 64;
        for (int i = 0; i < ns && !q.err; ++i)
        {
            const size_t t = (size_t)j * kSegSlots + i;
            int basel = segresume[t];
            if (basel < 0) continue;
            const int stop = i + 1 < ns ? st[i + 1] : -1;
            const int2 sp = segpool[t];
            Pool pool = { sp.x, sp.y };
            const int r = q.bubble(basel, stop, pool, budget);
            if (stop >= 0 && r != 1 && !q.err) q.err = 4;
            segpool[t] = make_int2(pool.head, pool.tail);
        }
        resumed = (unsigned long long)nstop[ln];
    }
    wave_add(&counters[7], ran);
    wave_add(&counters[8], resumed);
    wave_add(&counters[9], stopped);
    flush_counts(counters, q.pc);
    if (q.err) atomicOr(err, q.err >= 4 ? kSeamErr : (2 << q.err));
}

// one workgroup per node: the free lists of the node, its pieces and its
// anchors are joined in parallel (compacted by a block scan, each list's
// tail linked to the next list's head); then the hull re-walk
__global__ __launch_bounds__(kSegThreads) void k_seg_final(int* NX, int* OR, const double2* PT, int* VE,
                                                       const int* off, int d, int K, uint32_t m, int lim_e,
                                                       const int* nseg, const int4* nstate, const int2* segpool,
                                                       const int* astate, const int2* apool, int4* res, int* err,
                                                       unsigned long long* counters)
{
    // lists per thread: 1 + kSegSlots + kSegMaxK of them at most
    constexpr int kPer = (1 + kSegSlots + kSegMaxK + kSegThreads - 1) / kSegThreads;
    __shared__ int2 lists[kPer * kSegThreads];
    const uint32_t j = blockIdx.x;
    const int k = threadIdx.x;
    if (j >= lcount(off, d)) return;
    const int* od = off + off_base(off, d);
    const int s = od[j + 1] - od[j];
    if (s < 4) return;
    const int4 ns = nstate[j];
    const int cnt = nseg[j];
    int2 mine[kPer];
    uint32_t c = 0;
#pragma unroll
    for (int r = 0; r < kPer; ++r)
    {
        const int x = kPer * k + r;
        int2 l = make_int2(-1, -1);
        if (x == 0)
            l = make_int2(ns.z, ns.w);
        else if (x <= cnt)
            l = segpool[(size_t)j * kSegSlots + (x - 1)];
        else if (x <= cnt + K)
        {
            const size_t t = (size_t)j * K + (x - cnt - 1);
            if (astate[t] == 2) l = apool[t];
        }
        mine[r] = l;
        c += l.x >= 0 ? 1u : 0u;
    }
    uint32_t total;
    uint32_t pos = block_exclusive_scan(c, total);
#pragma unroll
    for (int r = 0; r < kPer; ++r)
        if (mine[r].x >= 0) lists[pos++] = mine[r];
    __syncthreads();
    const int lim_r = lim_e >> 2;
    for (uint32_t x = k; x + 1 < total; x += kSegThreads)
    {
        const int tail = lists[x].y;
        if ((unsigned)tail < (unsigned)lim_r)
            NX[4 * tail] = lists[x + 1].x;
        else
            atomicOr(err, 2 << 1);
    }
    if (k != 0) return;
    const int head = total ? lists[0].x : -1, tail = total ? lists[total - 1].y : -1;
    // the hull edges in this level's frame; below the root the parent's frame
    // differs and its first reader re-walks them (child_result)
    const bool rewalk = d > 0 && ns.x >= 0;
    res[node_index(off, d, j)] = make_int4(rewalk ? -(ns.x + 2) : ns.x, ns.y, head, tail);
}


// ---------------------------------------------------------------------------
// 8: extraction
// ---------------------------------------------------------------------------

// face[k]: bit 0/1 = the face left of edge 4k / 4k+2 is an output triangle,
// bit 2/3 = undecided by the filter (queued)
// Face of directed edge e: its lnext 3-cycle (f, g), or false.  Indices are
// checked so that a damaged store cannot send a load out of bounds.
__device__ __forceinline__ bool face_of(const int* NX, const int* OR, int e, int lim_e, int lim_v, int& f, int& g,
                                        int& a, int& b, int& c)
{
    f = e_rot(NX[e_invrot(e)]);
    if ((unsigned)f >= (unsigned)lim_e) return false;
    g = e_rot(NX[e_invrot(f)]);
    if ((unsigned)g >= (unsigned)lim_e) return false;
    if (e_rot(NX[e_invrot(g)]) != e) return false;
    a = OR[e >> 1];
    b = OR[f >> 1];
    c = OR[g >> 1];
    return (unsigned)a < (unsigned)lim_v && (unsigned)b < (unsigned)lim_v && (unsigned)c < (unsigned)lim_v;
}

// The left faces of edges 4k and 4k+2 of record k (its two primal edges),
// walked side by side so that each step's loads for both faces are in flight
// together: ok[i] when the lnext cycle of edge 4k+2i closes after three
// edges within bounds (and, with `smallest`, when no edge of it is below the
// starting one, so each face is seen from one edge only); v[i] its origins.
__device__ __forceinline__ void record_faces(const int* NX, const int* OR, uint32_t k, int lim_e, int lim_v,
                                             bool smallest, uint32_t want, bool ok[2], int v[2][3])
{
    const int4 R = reinterpret_cast<const int4*>(NX)[k]; // onext of 4k .. 4k+3
    const int e[2] = { 4 * (int)k, 4 * (int)k + 2 };
    int f[2], g[2];
    f[0] = e_rot(R.w); // lnext(4k) = rot(onext(invrot(4k))) = rot(onext(4k+3))
    f[1] = e_rot(R.y);
#pragma unroll
    for (int i = 0; i < 2; ++i)
        ok[i] = (want >> i & 1u) && (unsigned)f[i] < (unsigned)lim_e && (!smallest || f[i] >= e[i]);
    int t[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) t[i] = ok[i] ? NX[e_invrot(f[i])] : 0;
#pragma unroll
    for (int i = 0; i < 2; ++i)
    {
        g[i] = e_rot(t[i]);
        ok[i] = ok[i] && (unsigned)g[i] < (unsigned)lim_e && (!smallest || g[i] >= e[i]);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) t[i] = ok[i] ? NX[e_invrot(g[i])] : 0;
#pragma unroll
    for (int i = 0; i < 2; ++i)
    {
        ok[i] = ok[i] && e_rot(t[i]) == e[i];
        v[i][0] = ok[i] ? OR[e[i] >> 1] : 0;
        v[i][1] = ok[i] ? OR[f[i] >> 1] : 0;
        v[i][2] = ok[i] ? OR[g[i] >> 1] : 0;
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
        ok[i] = ok[i] && (unsigned)v[i][0] < (unsigned)lim_v && (unsigned)v[i][1] < (unsigned)lim_v &&
                (unsigned)v[i][2] < (unsigned)lim_v;
}

__global__ __launch_bounds__(kBlock) void k_extract_test(const int* NX, const int* OR, const double2* PT,
                                                         uint32_t nrec, uint32_t m, int lim_e, const unsigned* spare,
                                                         uint32_t spare_cap, uint32_t* face, uint32_t* queue,
                                                         uint32_t qcap, unsigned long long* counters, int* err)
{
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= nrec) return;
    const uint32_t used = 3 * m + min(*spare, spare_cap);
    uint32_t fl = 0;
    if (k < used && OR[2 * k] >= 0)
    {
        // the face is reported by its smallest edge: the walk stops at the
        // first smaller edge of the cycle before reading further
        bool ok[2];
        int v[2][3];
        record_faces(NX, OR, k, lim_e, (int)m, true, 3u, ok, v);
        double2 PA[2], PB[2], PC[2];
#pragma unroll
        for (int i = 0; i < 2; ++i)
            if (ok[i])
            {
                PA[i] = PT[v[i][0]];
                PB[i] = PT[v[i][1]];
                PC[i] = PT[v[i][2]];
            }
#pragma unroll
        for (int r = 0; r < 4; r += 2)
        {
            if (!ok[r >> 1]) continue;
            const int e = 4 * (int)k + r;
            const double2 A = PA[r >> 1], B = PB[r >> 1], C = PC[r >> 1];
            const double detleft = (A.x - C.x) * (B.y - C.y);
            const double detright = (A.y - C.y) * (B.x - C.x);
            const double det = detleft - detright;
            const double mag = fabs(detleft) + fabs(detright);
            const double bound = kErrOrient * mag;
            if ((det > bound || -det > bound) && mag > kFilterFloor)
            {
                if (det > 0) fl |= 1u << (r >> 1);
            }
            else
            {
                // queued for k_extract_exact; past the queue's end the face
                // is marked for k_extract_overflow instead
                const unsigned long long slot = atomicAdd(&counters[4], 1ull);
                if (slot < qcap)
                {
                    fl |= 4u << (r >> 1);
                    queue[slot] = (uint32_t)e;
                }
                else
                    fl |= 16u << (r >> 1);
            }
        }
    }
    face[k] = fl;
}

__global__ __launch_bounds__(kBlock) void k_extract_exact(const int* NX, const int* OR, const double2* PT,
                                                          uint32_t m, int lim_e, const uint32_t* queue,
                                                          const unsigned long long* counters, uint32_t qcap,
                                                          uint32_t* face)
{
    const unsigned long long cnt = counters[4] < qcap ? counters[4] : qcap;
    for (unsigned long long t = blockIdx.x * kBlock + threadIdx.x; t < cnt; t += (unsigned long long)gridDim.x * kBlock)
    {
        const int e = (int)queue[t];
        int f, g, a, b, c;
        if (!face_of(NX, OR, e, lim_e, (int)m, f, g, a, b, c)) continue;
        const double2 A = PT[a], B = PT[b], C = PT[c];
        if (orient2d_exact(A.x, A.y, B.x, B.y, C.x, C.y) > 0) atomicOr(&face[e >> 2], 1u << ((e & 2) >> 1));
    }
}

// Faces the queue had no room for (inputs whose every face escapes the
// filter, e.g. all coordinates below ~1e-126 or above ~1e154, where the
// filter's products fall under its floor or overflow): decided exactly here,
// by a grid-stride pass over the records that ends at once when the queue
// did not overflow.
__global__ __launch_bounds__(kBlock) void k_extract_overflow(const int* NX, const int* OR, const double2* PT,
                                                             uint32_t nrec, uint32_t m, int lim_e,
                                                             const unsigned long long* counters, uint32_t qcap,
                                                             uint32_t* face)
{
    if (counters[4] <= qcap) return;
    for (uint32_t k = blockIdx.x * kBlock + threadIdx.x; k < nrec; k += gridDim.x * kBlock)
    {
        const uint32_t fl = face[k];
        if (!(fl & 48u)) continue;
        uint32_t add = 0;
        for (int i = 0; i < 2; ++i)
        {
            if (!(fl >> (4 + i) & 1u)) continue;
            const int e = 4 * (int)k + 2 * i;
            int f, g, a, b, c;
            if (!face_of(NX, OR, e, lim_e, (int)m, f, g, a, b, c)) continue;
            const double2 A = PT[a], B = PT[b], C = PT[c];
            if (orient2d_exact(A.x, A.y, B.x, B.y, C.x, C.y) > 0) add |= 1u << i;
        }
        face[k] = (fl & ~48u) | add;
    }
}

struct FaceCount
{
    const uint32_t* face;
    __device__ uint32_t operator()(uint32_t k) const { return __popc(face[k] & 3u); }
};

// the scan's output: exclusive triangle count of every record, packed above
// its two face bits
struct FacePrefix
{
    uint32_t* face;
    __device__ uint32_t load(uint32_t k) const { return face[k]; }
    __device__ void store(uint32_t k, uint32_t pre, uint32_t f) const { face[k] = (pre << 2) | (f & 3u); }
};

// one thread per record: writes its faces at their prefix (as original,
// lowest-duplicate indices)
__global__ __launch_bounds__(kBlock) void k_face_write(const uint32_t* face, const int* NX, const int* OR,
                                                       const uint32_t* ID0, uint32_t nrec, uint32_t m, int lim_e,
                                                       uint32_t split, uint32_t* tmain, uint32_t* tris,
                                                       uint32_t tri_cap, int* err)
{
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= nrec) return;
    const uint32_t f = face[k];
    uint32_t pre = f >> 2;
    if (k == split) *tmain = pre; // triangles of the records below the spare region
    const uint32_t fl = f & 3u;
    if (!fl) return;
    // a store that is not a planar triangulation can show more faces than a
    // triangulation of m points has (at most 2m - 5): never past the buffer
    if (pre + (uint32_t)__popc(fl) > tri_cap)
    {
        atomicOr(err, 1 << 22);
        return;
    }
    // both faces' loads before any store
    bool ok[2];
    int v[2][3];
    record_faces(NX, OR, k, lim_e, (int)m, false, fl, ok, v);
    uint32_t id[2][3];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) id[i][j] = (fl >> i & 1u) ? ID0[ok[i] ? v[i][j] : 0] : 0;
#pragma unroll
    for (int i = 0; i < 2; ++i)
    {
        if (!(fl >> i & 1u)) continue;
        tris[3 * pre] = id[i][0];
        tris[3 * pre + 1] = id[i][1];
        tris[3 * pre + 2] = id[i][2];
        ++pre;
    }
}

// ---------------------------------------------------------------------------
// 9: exact verifier of a triangle list
// ---------------------------------------------------------------------------

// verifier counters
enum : int
{
    kvBoundary,
    kvVertices,
    kvBadIndex,
    kvNotRep,
    kvBadOrient,
    kvBadEdges,
    kvNonDelaunay,
    kvConcave,
    kvUnused,
    kvSecond,  // lowest representative index with coordinates other than point 0's
    kvOffLine, // representatives off the line through point 0 and that one
    kvCount
};

constexpr uint64_t kNoEdge = ~0ull;

// per triangle: index checks, exact orientation, used vertices, and its
// three edges keyed by the unordered vertex pair (value 3t + i: edge i runs
// from corner i to corner i + 1)
__global__ __launch_bounds__(kBlock) void k_ver_tris(const double2* P, uint32_t n, const uint32_t* tris, uint32_t T,
                                                     const uint32_t* rep, uint64_t* keys, uint32_t* vals,
                                                     uint8_t* used, unsigned long long* cnt)
{
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= T) return;
    const uint32_t v[3] = { tris[3 * (size_t)t], tris[3 * (size_t)t + 1], tris[3 * (size_t)t + 2] };
    const bool idx_ok = v[0] < n && v[1] < n && v[2] < n && v[0] != v[1] && v[1] != v[2] && v[0] != v[2];
    if (!idx_ok)
    {
        atomicAdd(&cnt[kvBadIndex], 1ull);
        for (int i = 0; i < 3; ++i)
        {
            keys[3 * (size_t)t + i] = kNoEdge;
            vals[3 * (size_t)t + i] = 3 * t + i;
        }
        return;
    }
    if (rep && (rep[v[0]] != v[0] || rep[v[1]] != v[1] || rep[v[2]] != v[2])) atomicAdd(&cnt[kvNotRep], 1ull);
    const double2 A = P[v[0]], B = P[v[1]], C = P[v[2]];
    unsigned nex = 0;
    if (!(orient2d(A.x, A.y, B.x, B.y, C.x, C.y, nex) > 0)) atomicAdd(&cnt[kvBadOrient], 1ull);
    for (int i = 0; i < 3; ++i)
    {
        used[v[i]] = 1;
        const uint32_t a = v[i], b = v[(i + 1) % 3];
        keys[3 * (size_t)t + i] = ((uint64_t)min(a, b) << 32) | max(a, b);
        vals[3 * (size_t)t + i] = 3 * t + i;
    }
}

// per group of equal edge keys (sorted): one edge = boundary (its direction
// recorded as the boundary successor of its origin), two = interior (opposite
// directions, exact incircle of the far corner), more = not a manifold
__global__ __launch_bounds__(kBlock) void k_ver_edges(const double2* P, const uint32_t* tris, const uint64_t* keys,
                                                      const uint32_t* vals, uint32_t E, int* nextb,
                                                      unsigned long long* cnt)
{
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= E) return;
    const uint64_t key = keys[p];
    if (key == kNoEdge || (p > 0 && keys[p - 1] == key)) return;
    uint32_t g = 1;
    while (p + g < E && keys[p + g] == key && g < 3) ++g;
    auto corner = [&](uint32_t e, int k) { return tris[3 * (size_t)(e / 3) + (e % 3 + k) % 3]; };
    const uint32_t e1 = vals[p];
    const uint32_t u1 = corner(e1, 0), v1 = corner(e1, 1);
    if (g >= 3)
    {
        atomicAdd(&cnt[kvBadEdges], 1ull);
        return;
    }
    if (g == 1)
    {
        atomicAdd(&cnt[kvBoundary], 1ull);
        if (atomicCAS(&nextb[u1], -1, (int)v1) != -1) atomicAdd(&cnt[kvBadEdges], 1ull);
        return;
    }
    const uint32_t e2 = vals[p + 1];
    const uint32_t u2 = corner(e2, 0), v2 = corner(e2, 1);
    if (u1 != v2 || v1 != u2)
    {
        atomicAdd(&cnt[kvBadEdges], 1ull);
        return;
    }
    const double2 A = P[u1], B = P[v1], C = P[corner(e1, 2)], Dp = P[corner(e2, 2)];
    unsigned nex = 0;
    if (incircle(A.x, A.y, B.x, B.y, C.x, C.y, Dp.x, Dp.y, nex) > 0) atomicAdd(&cnt[kvNonDelaunay], 1ull);
}

// With no triangles the distinct points must be collinear: a second point
// (the lowest representative not equal to point 0; point 0 is always a
// representative) fixes the line, and every representative is tested.
__global__ __launch_bounds__(kBlock) void k_ver_second(const double2* P, uint32_t n, const uint32_t* rep,
                                                       unsigned long long* cnt)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n || (rep && rep[i] != i)) return;
    const double2 a = P[0], b = P[i];
    if (a.x != b.x || a.y != b.y) atomicMin(&cnt[kvSecond], (unsigned long long)i);
}

__global__ __launch_bounds__(kBlock) void k_ver_line(const double2* P, uint32_t n, const uint32_t* rep,
                                                     unsigned long long* cnt)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const unsigned long long s = cnt[kvSecond];
    if (i >= n || s >= n || (rep && rep[i] != i)) return;
    const double2 a = P[0], b = P[s], c = P[i];
    unsigned nex = 0;
    if (orient2d(a.x, a.y, b.x, b.y, c.x, c.y, nex) != 0) atomicAdd(&cnt[kvOffLine], 1ull);
}

// per vertex: its boundary turn (exact orient2d of predecessor edge and
// successor), used-vertex count, unused representatives
__global__ __launch_bounds__(kBlock) void k_ver_vertices(const double2* P, uint32_t n, const uint32_t* rep,
                                                         const uint8_t* used, const int* nextb,
                                                         unsigned long long* cnt)
{
    const uint32_t u = blockIdx.x * kBlock + threadIdx.x;
    bool is_used = false, unused_rep = false;
    if (u < n)
    {
        is_used = used[u] != 0;
        unused_rep = rep && !is_used && rep[u] == u;
        const int v = nextb[u];
        if (v >= 0)
        {
            const int w = nextb[v];
            if (w < 0)
                atomicAdd(&cnt[kvBadEdges], 1ull); // the boundary does not continue at v
            else
            {
                const double2 A = P[u], B = P[v], C = P[w];
                unsigned nex = 0;
                if (orient2d(A.x, A.y, B.x, B.y, C.x, C.y, nex) < 0) atomicAdd(&cnt[kvConcave], 1ull);
            }
        }
    }
    const unsigned long long nu = __popcll(__ballot(is_used)), nr = __popcll(__ballot(unused_rep));
    if (__lane_id() == 0)
    {
        if (nu) atomicAdd(&cnt[kvVertices], nu);
        if (nr) atomicAdd(&cnt[kvUnused], nr);
    }
}

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------

template <class In, class Out>
void scan_u32(In in, uint32_t n, Out out, uint32_t* partials, hipStream_t s)
{
    const uint32_t nb = cdiv(n, kScanTile);
    k_scan_reduce<<<nb, kBlock, 0, s>>>(in, n, partials);
    k_scan_partials<<<1, 1024, 0, s>>>(partials, nb);
    k_scan_down<<<nb, kBlock, 0, s>>>(in, n, partials, out);
}

// the same with separate inputs for the two passes (the first may leave its
// values where the second reads them)
template <class InReduce, class InDown, class Out>
void scan_u32(InReduce in1, InDown in2, uint32_t n, Out out, uint32_t* partials, hipStream_t s)
{
    const uint32_t nb = cdiv(n, kScanTile);
    k_scan_reduce<<<nb, kBlock, 0, s>>>(in1, n, partials);
    k_scan_partials<<<1, 1024, 0, s>>>(partials, nb);
    k_scan_down<<<nb, kBlock, 0, s>>>(in2, n, partials, out);
}

// Kd depth at which every node has at most three points.
int kd_depth(uint32_t m)
{
    std::vector<uint32_t> sizes{ m }, next;
    int d = 0;
    for (;;)
    {
        bool any = false;
        next.clear();
        for (uint32_t s : sizes)
        {
            if (s >= 4)
            {
                any = true;
                next.push_back(s / 2);
                next.push_back(s - s / 2);
            }
            else
                next.push_back(s);
        }
        if (!any) return d;
        std::sort(next.begin(), next.end());
        next.erase(std::unique(next.begin(), next.end()), next.end());
        sizes.swap(next);
        ++d;
    }
}

} // namespace

struct Engine::Bufs
{
    // point-proportional (ensure)
    uint64_t *kA, *kB, *kx, *ky;
    uint32_t *vA, *vB, *dist, *ID0, *rep, *B0, *B1, *Bt, *ID1, *Z, *tris, *face, *queue;
    uint32_t *partials, *hist, *ghist, *vrank, *rin, *one_set, *rblock;
    double2* P0;
    double2* P1; // the other order's coordinates (vertex ids renumbered for the cells)
    uint8_t* flagp;
    int *NX, *OR;
    unsigned long long* counters;
    int* err;
    int *VE, *owner;
    unsigned* spare;
    size_t rec_cap; // quad-edge records allocated (3N + spare)
    uint32_t qcap;
    // node-proportional (ensure_nodes)
    int* off; // level table, then every level's node offsets
    int4* res;
    int *splitv, *qv, *kn, *lvlK, *redo, *segstart, *nseg, *segresume;
    int2* ext; // per node: its extreme points in its parent's frame (min, max)
    int *arec, *astate, *aedge, *selk, *nsel, *alist, *alist_n;
    double4* acoord;
    double* akey;
    int2 *apool, *segpool;
    int4* nstate;
    size_t nodes_cap, offs_cap, seg_nodes, anchor_cap, qv_nodes;
    // per set (ensure_sets)
    uint32_t *soff, *srank, *rpos, *vstart, *setbad, *in_off, *mstart, *sstart, *toff, *trank, *btris, *tmain;
    size_t sets_cap, tris_cap;
    // pinned host staging
    uint32_t* host_scalars; // 64 words, pinned: scalars and the result block read back
    uint32_t* host_hist;    // 2 x 8 x 256 words, pinned: the sorts' digit histograms read back
    int* host_table;
    uint32_t* host_sets;
    size_t host_sets_cap;
};

// kd forest plan: every level's node count and the ranks rooted there
struct Engine::Plan
{
    int D = 0;                // leaf level
    std::vector<int> count;   // nodes per level
    std::vector<int> offb;    // first node offset of the level (after the table)
    std::vector<int> nodeb;   // first node index of the level
    std::vector<int> ra;      // ranks [ra[d], ra[d+1]) are rooted at level d
    std::vector<double> smax; // largest node of the level (bound)
    std::vector<int> K;       // anchors per node at the level (0: not cut into pieces)
    size_t nodes = 0, offs = 0, seg_nodes = 1, anchor_cap = 1;
    size_t qv_nodes = 1; // node index bound of the levels cut into pieces (qv is indexed by node)
};

// The values a batch run took host decisions on (recorded by a run in
// record mode, replayed by the captured graph) and the pinned memory its
// graph copies from.
struct Engine::Trace
{
    std::vector<uint32_t> sort_mask; // per sort of the run, in order: the radix passes run
    std::vector<uint32_t> setbad;    // per set: a non-finite point
    uint32_t m = 0;                  // distinct points
    std::vector<uint32_t> vstart;    // first distinct point of every rank (k + 1)
    bool ytie = false;               // a run of equal y too long for the in-place fix
    std::vector<int> K;              // anchors per node of every kd level after the kd phase
    uint32_t T = 0, tmain = 0;       // triangles of the recorded run, and those below the spare records
    uint32_t T_bound = 0, c_bound = 0; // replay bounds on the triangles and the spare records' triangles
    std::vector<int> set_rank;       // rank of every set in the kd forest
    int kg = 0;                      // sets in the forest
    size_t upload_bytes = 0;         // host-to-device bytes of a run (staging size)
    char* stage = nullptr;           // pinned staging of a replay's uploads (owned)
    size_t stage_cap = 0, stage_used = 0;
    ~Trace()
    {
        if (stage) (void)hipHostFree(stage);
    }
};

struct Engine::GraphEntry
{
    std::vector<size_t> sizes; // the shape: points per set
    bool seams = true;
    int cell_max = 0, merge_lanes = 0;
    uint64_t ws_gen = 0;
    Trace tr;
    // two graphs: the first runs up to the distinct points and ends with
    // the guards on every count the rest of the launches are sized by; the
    // second (the kd forest, merges, extraction, output) is launched only
    // when those guards hold
    hipGraph_t graph = nullptr, graph2 = nullptr;
    hipGraphExec_t exec = nullptr, exec2 = nullptr;
    hipGraphNode_t in_node = nullptr; // the copy of the caller's points (its source is set per launch)
    const double* in_src = nullptr;
    uint32_t* rb = nullptr;           // pinned: the result block and the ranks' triangle offsets
    uint64_t uses = 0;
    ~GraphEntry()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (exec2) (void)hipGraphExecDestroy(exec2);
        if (graph) (void)hipGraphDestroy(graph);
        if (graph2) (void)hipGraphDestroy(graph2);
        if (rb) (void)hipHostFree(rb);
    }
};

namespace
{

// sizes[r], roots[r]: size bound and root level of the set of rank r (roots
// ascending); D: leaf level
void make_plan(Engine::Plan& P, const std::vector<uint32_t>& sizes, const std::vector<int>& roots, int D, bool seams)
{
    P.D = D;
    P.count.assign(D + 1, 0);
    P.offb.assign(D + 1, 0);
    P.nodeb.assign(D + 1, 0);
    P.ra.assign(D + 2, (int)sizes.size());
    P.smax.assign(D + 1, 0.0);
    P.K.assign(D + 1, 0);
    for (int r = (int)sizes.size() - 1; r >= 0; --r) P.ra[roots[r]] = r;
    for (int d = D; d >= 0; --d) P.ra[d] = std::min(P.ra[d], P.ra[d + 1]);
    size_t offs = 0, nodes = 0;
    for (int d = 0; d <= D; ++d)
    {
        P.count[d] = (d ? 2 * P.count[d - 1] : 0) + (P.ra[d + 1] - P.ra[d]);
        P.offb[d] = (int)offs;
        P.nodeb[d] = (int)nodes;
        offs += (size_t)P.count[d] + 1;
        nodes += (size_t)P.count[d];
    }
    for (size_t r = 0; r < sizes.size(); ++r)
    {
        double sz = sizes[r];
        for (int d = roots[r]; d <= D; ++d)
        {
            P.smax[d] = std::max(P.smax[d], sz);
            if (sz >= 4) sz = std::ceil(sz / 2);
        }
    }
    P.nodes = nodes;
    P.offs = offs;
    P.seg_nodes = 1;
    P.anchor_cap = 1;
    P.qv_nodes = 1;
    for (int d = 0; d < D; ++d)
    {
        // candidate levels (the nodes report their own anchor counts during
        // the kd phase)
        P.K[d] = seams ? seam_anchors_max(P.smax[d]) : 0;
        if (!P.K[d]) continue;
        P.seg_nodes = std::max(P.seg_nodes, (size_t)P.count[d]);
        P.qv_nodes = std::max(P.qv_nodes, (size_t)P.nodeb[d] + (size_t)P.count[d]);
        P.anchor_cap = std::max(P.anchor_cap, (size_t)P.count[d] * P.K[d]);
    }
}

constexpr size_t kHostSetArrays = 8;

} // namespace

struct Item
{
    void** p;
    size_t bytes;
};

// One device region holding the items (each 256-byte aligned): a fresh
// allocation, or the next part of the external workspace.
template <size_t N>
Status Engine::carve_region(const Item (&items)[N], char*& region, size_t& bytes)
{
    size_t total = 0;
    for (const Item& it : items) total += (it.bytes + 255) & ~(size_t)255;
    if (mode_ == Mode::replay) return Status::internal; // a replay reuses the recorded run's regions
    ++ws_gen_;
    if (ext_)
    {
        if (total > ext_bytes_ - ext_used_) return Status::out_of_memory;
        region = ext_ + ext_used_;
        ext_used_ += total;
    }
    else
    {
        if (total > alloc_budget()) return Status::out_of_memory;
        void* ws = nullptr;
        const hipError_t e = hipMalloc(&ws, total);
        if (e != hipSuccess) return e == hipErrorOutOfMemory ? Status::out_of_memory : Status::internal;
        region = (char*)ws;
    }
    bytes = total;
    size_t at = 0;
    for (const Item& it : items)
    {
        *it.p = region + at;
        at += (it.bytes + 255) & ~(size_t)255;
    }
    return Status::ok;
}

void Engine::free_region(char*& region, size_t& bytes)
{
    if (region && !ext_) (void)hipFree(region);
    region = nullptr;
    bytes = 0;
}

Engine::Engine(hipStream_t stream, void* ext, size_t ext_bytes)
    : stream_(stream), own_stream_(false), ext_((char*)ext), ext_bytes_(ext_bytes)
{
    b_ = new Bufs();
    std::memset(b_, 0, sizeof(Bufs));
    b_->host_scalars = ext_scalars_;
    b_->host_table = ext_table_;
    b_->host_hist = ext_hist_;
}

Engine::Engine(hipStream_t stream) : stream_(stream), own_stream_(stream == nullptr)
{
    if (own_stream_)
    {
        if (hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking) != hipSuccess) stream_ = nullptr;
    }
    b_ = new Bufs();
    std::memset(b_, 0, sizeof(Bufs));
}

hipEvent_t Engine::prof_event(size_t i)
{
    while (evs_.size() <= i)
    {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        evs_.push_back(e);
    }
    return evs_[i];
}

Engine::~Engine()
{
    for (hipEvent_t e : evs_) (void)hipEventDestroy(e);
    release();
    if (b_ && !ext_)
    {
        if (b_->host_scalars) (void)hipHostFree(b_->host_scalars);
        if (b_->host_table) (void)hipHostFree(b_->host_table);
        if (b_->host_hist) (void)hipHostFree(b_->host_hist);
        if (b_->host_sets) (void)hipHostFree(b_->host_sets);
    }
    delete b_;
    if (ev0_) (void)hipEventDestroy(ev0_);
    if (ev1_) (void)hipEventDestroy(ev1_);
    if (own_stream_ && stream_) (void)hipStreamDestroy(stream_);
}

void Engine::release()
{
    free_region(ws_, ws_bytes_);
    free_region(ws_nodes_, ws_nodes_bytes_);
    free_region(ws_anchors_, ws_anchors_bytes_);
    free_region(ws_sets_, ws_sets_bytes_);
    free_region(ws_gin_, ws_gin_bytes_);
    gin_ = nullptr;
    gin_cap_ = 0;
    drop_graphs(); // they hold the regions' addresses
    ext_used_ = ext_ws_end_ = 0;
    cap_n_ = 0;
    b_->nodes_cap = b_->offs_cap = b_->seg_nodes = b_->anchor_cap = b_->qv_nodes = 0;
    b_->sets_cap = b_->tris_cap = 0;
    res_ = DeviceResult();
    batch_ = BatchResult();
}

size_t Engine::workspace_bytes() const
{
    return ws_bytes_ + ws_nodes_bytes_ + ws_anchors_bytes_ + ws_sets_bytes_ + ws_gin_bytes_;
}

namespace
{
size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }
} // namespace

// Device bytes the workspace will hold after a run over n points in k sets:
// every region keeps its size or grows to what the run needs (ensure,
// ensure_sets: the same sizes as there; the node and seam regions depend on
// the distinct points' tree, bounded here by 96 B per point and 8 MB).
size_t Engine::bytes_after(size_t n, size_t k) const
{
    const size_t N = std::max<size_t>(n, 64);
    size_t ws = ws_bytes_;
    if (n > cap_n_ || !ws_)
    {
        const size_t R = 3 * N + N / 2 + 65536;
        const size_t nb_scan = cdiv(R, kScanTile) + 1;
        const size_t qcap = std::max<size_t>(4096, N / 4);
        const size_t per[] = { 8 * N, 8 * N, 8 * N, 8 * N, 4 * N, 4 * N, 4 * N, 4 * N, 4 * N, 4 * N, 4 * N,
                               4 * N, 4 * N, 4 * N, 4 * 6 * N, 4 * R, 4 * qcap,
                               4 * (nb_scan + 1), 4 * radix_state_words(N), 4 * 2 * 8 * 256, 4 * N, 4 * N,
                               16 * N, 16 * N, N, 4 * 4 * R, 4 * 2 * R, 8 * 16, 64, 4 * N, 4 * N, 64, 64, 256 };
        ws = 0;
        for (size_t b : per) ws += round256(b);
    }
    size_t sets = ws_sets_bytes_;
    if (k)
    {
        const size_t tris = 2 * N;
        if (!ws_sets_ || k > b_->sets_cap || tris > b_->tris_cap)
        {
            const size_t K = std::max(k, b_->sets_cap) + 1;
            const size_t T = std::max(tris, b_->tris_cap);
            sets = 9 * round256(4 * K) + round256(4 * T) + round256(12 * T) + round256(64);
        }
    }
    const size_t tree = std::max(ws_nodes_bytes_ + ws_anchors_bytes_, 96 * N + (8u << 20));
    return std::max(ws, ws_bytes_) + std::max(sets, ws_sets_bytes_) + tree;
}

// bytes a new region may take under the test limit (regions being replaced
// are freed before they are carved again)
size_t Engine::alloc_budget() const
{
    if (!mem_cap_) return SIZE_MAX;
    const size_t held = workspace_bytes();
    return mem_cap_ > held ? mem_cap_ - held : 0;
}

constexpr uint32_t kSortErr = 1u << 23; // err bit: a radix pass's look-back gave up
// The two sorts of a single set run eagerly, from kSplitMin keys on, take
// radix passes over the top kSplitPasses bytes in which the keys differ and
// finish with k_sort_fixup (10M uniform points: 5 passes and a fix-up
// instead of 8 passes per sort; the runs of keys equal in those bytes hold
// ~1.02 keys on average).  A long run out of order sets kSplitErr, and the
// run is repeated with full sorts.
constexpr uint32_t kSplitErr = 1u << 22;
constexpr int kSplitPasses = 5;
constexpr uint32_t kSplitMin = 1u << 20;

// The radix passes of a sort (bit p of mask: pass p runs); the result ends
// in kA/vA.  ghist: the keys' digit histograms (k_radix_global_hist); state:
// radix_state_words(n) words of look-back state; a look-back that gives up
// sets err_bit in *err.
static void sort_passes(uint64_t*& kA, uint32_t*& vA, uint64_t*& kB, uint32_t*& vB, uint32_t n, uint32_t mask,
                        const uint32_t* ghist, uint32_t* state, uint32_t* err, uint32_t err_bit, hipStream_t s)
{
    if (!mask) return;
    const uint32_t nb = cdiv(n, kSortTile);
    const size_t half = radix_half_words(n);
    // the first pass's half; each pass clears the other for the next
    fill_words(state, half, 0, s);
    int h = 0;
    for (int pass = 0; pass < 8; ++pass)
    {
        if (!(mask >> pass & 1u)) continue;
        k_radix_pass<<<nb, kBlock, 0, s>>>(kA, vA, kB, vB, n, 8 * pass, ghist, state + h * half,
                                           state + (1 - h) * half, err, err_bit);
        h ^= 1;
        std::swap(kA, kB);
        std::swap(vA, vB);
    }
}

// the passes whose digit is not the same for all n keys
static uint32_t pass_mask(const uint32_t* host_ghist, uint32_t n)
{
    uint32_t mask = 0;
    for (int pass = 0; pass < 8; ++pass)
    {
        bool trivial = false;
        for (int dgt = 0; dgt < 256; ++dgt)
            if (host_ghist[pass * 256 + dgt] == n) trivial = true;
        if (!trivial) mask |= 1u << pass;
    }
    return mask;
}

// the radix passes that can vary among keys in [0, v]
static uint32_t passes_upto(uint64_t v)
{
    uint32_t mask = 0;
    for (int p = 0; p < 8 && (v >> (8 * p)); ++p) mask |= 1u << p;
    return mask;
}

// guard of a replayed sort: every pass the recorded mask skips must still
// have one digit for all n keys
__global__ __launch_bounds__(kBlock) void k_guard_sort(const uint32_t* ghist, uint32_t n, uint32_t mask,
                                                       unsigned long long* guard)
{
    __shared__ int trivial[8];
    if (threadIdx.x < 8) trivial[threadIdx.x] = 0;
    __syncthreads();
    for (int pass = 0; pass < 8; ++pass)
        if (ghist[pass * 256 + threadIdx.x] == n) trivial[pass] = 1;
    __syncthreads();
    if (threadIdx.x < 8 && !(mask >> threadIdx.x & 1u) && !trivial[threadIdx.x]) atomicOr(guard, 1ull);
}

// guards of a replay: a device value against the recorded one
__global__ __launch_bounds__(kBlock) void k_guard_words(const uint32_t* got, const uint32_t* want, uint32_t n,
                                                        unsigned long long* guard, unsigned long long bit)
{
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
        if (got[i] != want[i]) atomicOr(guard, bit);
}

__global__ void k_guard_scalar(const uint32_t* got, uint32_t want, unsigned long long* guard, unsigned long long bit)
{
    if (threadIdx.x == 0 && *got != want) atomicOr(guard, bit);
}

__global__ void k_guard_le(const uint32_t* got, uint32_t bound, unsigned long long* guard, unsigned long long bit)
{
    if (threadIdx.x == 0 && *got > bound) atomicOr(guard, bit);
}

__global__ void k_guard_flag(const unsigned long long* got, int want_nonzero, unsigned long long* guard,
                             unsigned long long bit)
{
    if (threadIdx.x == 0 && (*got != 0) != (want_nonzero != 0)) atomicOr(guard, bit);
}

// Sorts (k, v) pairs stably by key; the result ends in kA/vA.  state:
// radix_state_words(n) words (its second word reports a look-back that gave
// up, read back here).
static Status sort_pairs(uint64_t*& kA, uint32_t*& vA, uint64_t*& kB, uint32_t*& vB, uint32_t n, uint32_t* state,
                         uint32_t* ghist, uint32_t* host_ghist, hipStream_t s)
{
    if (n <= 1) return Status::ok;
    if (n > kSortMaxKeys) return Status::internal;
    fill_words(ghist, 8 * 256, 0, s);
    k_radix_global_hist<<<std::min<uint32_t>(cdiv(n, kBlock), 1024), kBlock, 0, s>>>(kA, n, ghist);
    QDT_HIP(hipMemcpyAsync(host_ghist, ghist, 4 * 8 * 256, hipMemcpyDeviceToHost, s));
    QDT_HIP(hipStreamSynchronize(s));
    const uint32_t mask = pass_mask(host_ghist, n);
    if (!mask) return Status::ok;
    sort_passes(kA, vA, kB, vB, n, mask, ghist, state, state + 1, 1u, s);
    uint32_t gave_up = 0;
    QDT_HIP(hipMemcpyAsync(&gave_up, state + 1, 4, hipMemcpyDeviceToHost, s));
    QDT_HIP(hipStreamSynchronize(s));
    return gave_up ? Status::internal : Status::ok;
}

void Engine::set_tuning(int cell_points, int merge_lanes)
{
    const int c = cell_points <= 128 ? 128 : (cell_points <= 256 ? 256 : 512);
    const int l = merge_lanes <= 16 ? 16 : (merge_lanes <= 32 ? 32 : 64);
    if (c != cell_max_ || l != merge_lanes_) drop_graphs();
    cell_max_ = c;
    merge_lanes_ = l;
}

void Engine::drop_graphs()
{
    for (GraphEntry* g : graphs_cache_) delete g;
    graphs_cache_.clear();
}

// A host-to-device copy of data the run fixed on the host.  In replay mode
// (under capture) the bytes are kept in the graph's pinned staging, which
// the captured copy reads at every launch.
Status Engine::upload(void* dst, const void* src, size_t bytes)
{
    if (!bytes) return Status::ok;
    if (mode_ == Mode::replay)
    {
        Trace& t = *tr_;
        const size_t at = (t.stage_used + 255) & ~(size_t)255;
        if (at + bytes > t.stage_cap) return Status::internal;
        std::memcpy(t.stage + at, src, bytes);
        t.stage_used = at + bytes;
        QDT_HIP(hipMemcpyAsync(dst, t.stage + at, bytes, hipMemcpyHostToDevice, stream_));
        return Status::ok;
    }
    if (mode_ == Mode::record) tr_->upload_bytes += ((bytes + 255) & ~(size_t)255) + 256;
    QDT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream_));
    return Status::ok;
}

// Sorts (k, v) pairs stably by key (result in kA/vA), skipping the radix
// passes whose digit is the same for every key: read back in eager and
// record mode (and recorded), taken from the trace in replay mode with a
// device guard on the global histogram.
Status Engine::sort_keys(uint64_t*& kA, uint32_t*& vA, uint64_t*& kB, uint32_t*& vB, uint32_t n, uint32_t must)
{
    if (n <= 1) return Status::ok;
    if (n > kSortMaxKeys) return Status::internal;
    Bufs& b = *b_;
    hipStream_t s = stream_;
    fill_words(b.ghist, 8 * 256, 0, s);
    k_radix_global_hist<<<std::min<uint32_t>(cdiv(n, kBlock), 1024), kBlock, 0, s>>>(kA, n, b.ghist);
    uint32_t mask;
    if (mode_ == Mode::replay)
    {
        if (tr_cursor_ >= tr_->sort_mask.size()) return Status::internal;
        // a replay also runs the passes `must` names: every pass whose digit
        // can vary for this sort's keys, where a skipped pass would leave
        // them unsorted in launches the guards do not precede
        mask = tr_->sort_mask[tr_cursor_++] | must;
        k_guard_sort<<<1, 256, 0, s>>>(b.ghist, n, mask, &b.counters[15]);
    }
    else
    {
        QDT_HIP(hipMemcpyAsync(b.host_hist, b.ghist, 4 * 8 * 256, hipMemcpyDeviceToHost, s));
        QDT_HIP(hipStreamSynchronize(s));
        mask = pass_mask(b.host_hist, n);
        if (mode_ == Mode::record) tr_->sort_mask.push_back(mask);
    }
    sort_passes(kA, vA, kB, vB, n, mask, b.ghist, b.hist, (uint32_t*)b.err, kSortErr, s);
    return Status::ok;
}

Status Engine::ensure(size_t n)
{
    if (!stream_ && !ext_) return Status::internal;
    if (ext_)
    {
        // every run carves its regions afresh from the start of the workspace
        release();
        const Status st = ensure_ws(n);
        ext_ws_end_ = ext_used_;
        return st;
    }
    if (!b_->host_scalars)
    {
        void* h = nullptr;
        QDT_HIP(hipHostMalloc(&h, 4 * 64, hipHostMallocDefault));
        b_->host_scalars = (uint32_t*)h;
        QDT_HIP(hipHostMalloc(&h, 4 * kLvlHdr, hipHostMallocDefault));
        b_->host_table = (int*)h;
        QDT_HIP(hipHostMalloc(&h, 4 * 2 * 8 * 256, hipHostMallocDefault));
        b_->host_hist = (uint32_t*)h;
    }
    if (!ev0_)
    {
        QDT_HIP(hipEventCreate(&ev0_));
        QDT_HIP(hipEventCreate(&ev1_));
    }
    if (n <= cap_n_ && ws_) return Status::ok;
    return ensure_ws(n);
}

Status Engine::ensure_ws(size_t n)
{
    free_region(ws_, ws_bytes_);
    cap_n_ = 0;
    const size_t N = std::max<size_t>(n, 64);
    // Seam pieces start with empty pools: each takes the records for the
    // edges it adds beyond those it deletes (about the children's hull
    // vertices near its cut) from here, while as many records stay free in
    // the node's own region.  Summed over the seam levels this is ~n/8.
    const size_t spare_recs = N / 2 + 65536;
    const size_t R = 3 * N + spare_recs;
    const uint32_t nb_scan = cdiv(R, kScanTile) + 1;
    const size_t qcap = std::max<size_t>(4096, N / 4);
    Bufs& b = *b_;
    const Item items[] = {
        { (void**)&b.kA, 8 * N },       { (void**)&b.kB, 8 * N },       { (void**)&b.kx, 8 * N },
        { (void**)&b.ky, 8 * N },       { (void**)&b.vA, 4 * N },       { (void**)&b.vB, 4 * N },
        { (void**)&b.dist, 4 * N },     { (void**)&b.ID0, 4 * N },      { (void**)&b.rep, 4 * N },
        { (void**)&b.B0, 4 * N },       { (void**)&b.B1, 4 * N },       { (void**)&b.Bt, 4 * N },
        { (void**)&b.ID1, 4 * N },    { (void**)&b.Z, 4 * N },
        { (void**)&b.tris, 4 * 6 * N }, { (void**)&b.face, 4 * R },     { (void**)&b.queue, 4 * qcap },
        { (void**)&b.partials, 4 * (size_t)(nb_scan + 1) },
        { (void**)&b.hist, 4 * radix_state_words(N) },
        { (void**)&b.ghist, 4 * 2 * 8 * 256 },
        { (void**)&b.vrank, 4 * N },    { (void**)&b.rin, 4 * N },
        { (void**)&b.P0, 16 * N },      { (void**)&b.P1, 16 * N },      { (void**)&b.flagp, N },         { (void**)&b.NX, 4 * 4 * R },
        { (void**)&b.OR, 4 * 2 * R },   { (void**)&b.counters, 8 * 16 }, { (void**)&b.err, 64 },
        { (void**)&b.VE, 4 * N },       { (void**)&b.owner, 4 * N },
        { (void**)&b.spare, 64 },       { (void**)&b.one_set, 64 },     { (void**)&b.rblock, 256 },
    };
    const Status st = carve_region(items, ws_, ws_bytes_);
    if (st != Status::ok) return st;
    b.qcap = (uint32_t)qcap;
    b.rec_cap = R;
    cap_n_ = N;
    return Status::ok;
}

Status Engine::ensure_nodes(const Plan& P)
{
    Bufs& b = *b_;
    if (ws_nodes_ && P.nodes <= b.nodes_cap && P.offs <= b.offs_cap && P.qv_nodes <= b.qv_nodes) return Status::ok;
    free_region(ws_nodes_, ws_nodes_bytes_);
    // grow with headroom so that similar batches reuse the allocation
    const size_t nodes = std::max(P.nodes, b.nodes_cap) + 64;
    const size_t offs = std::max(P.offs, b.offs_cap) + 64;
    const size_t qv_nodes = std::max(P.qv_nodes, b.qv_nodes);
    const Item items[] = {
        { (void**)&b.off, 4 * (kLvlHdr + offs) }, { (void**)&b.res, 16 * nodes }, { (void**)&b.splitv, 4 * nodes },
        { (void**)&b.kn, 4 * nodes }, { (void**)&b.lvlK, 4 * kMaxLevels }, { (void**)&b.redo, 4 * (nodes + 1) },
        { (void**)&b.qv, 4 * (size_t)kSegMaxK * qv_nodes }, { (void**)&b.ext, 8 * nodes },
    };
    const Status st = carve_region(items, ws_nodes_, ws_nodes_bytes_);
    if (st != Status::ok) return st;
    b.nodes_cap = nodes;
    b.offs_cap = offs;
    b.qv_nodes = qv_nodes;
    return Status::ok;
}

// buffers of the seam pieces, sized by the levels the kd phase marked
Status Engine::ensure_anchors(size_t seg_nodes_need, size_t anchor_need)
{
    Bufs& b = *b_;
    if (ws_anchors_ && seg_nodes_need <= b.seg_nodes && anchor_need <= b.anchor_cap) return Status::ok;
    free_region(ws_anchors_, ws_anchors_bytes_);
    const size_t seg_nodes = std::max(seg_nodes_need, b.seg_nodes) + 1;
    const size_t anchor_cap = std::max(anchor_need, b.anchor_cap) + 1;
    const Item items[] = {
        { (void**)&b.segstart, 4 * (size_t)kSegSlots * seg_nodes },
        { (void**)&b.nseg, 4 * seg_nodes }, { (void**)&b.nstate, 16 * seg_nodes },
        { (void**)&b.segpool, 8 * (size_t)kSegSlots * seg_nodes },
        { (void**)&b.segresume, 4 * (size_t)kSegSlots * seg_nodes },
        { (void**)&b.arec, 4 * (size_t)kRec * anchor_cap }, { (void**)&b.acoord, 32 * anchor_cap },
        { (void**)&b.akey, 8 * anchor_cap },
        { (void**)&b.astate, 4 * anchor_cap }, { (void**)&b.aedge, 4 * anchor_cap },
        { (void**)&b.apool, 8 * anchor_cap }, { (void**)&b.selk, 4 * (size_t)kSegSlots * seg_nodes },
        { (void**)&b.nsel, 4 * seg_nodes },
        { (void**)&b.alist, 4 * (2 + 4 * (size_t)kCross) * anchor_cap }, { (void**)&b.alist_n, 64 },
    };
    const Status st = carve_region(items, ws_anchors_, ws_anchors_bytes_);
    if (st != Status::ok) return st;
    b.seg_nodes = seg_nodes;
    b.anchor_cap = anchor_cap;
    return Status::ok;
}

Status Engine::ensure_sets(size_t k, size_t n)
{
    Bufs& b = *b_;
    const size_t tris = 2 * std::max<size_t>(n, 64);
    if (!b.host_sets || b.host_sets_cap < k + 1)
    {
        if (b.host_sets) (void)hipHostFree(b.host_sets);
        b.host_sets = nullptr;
        void* h = nullptr;
        // staging for up to kHostSetArrays arrays of k + 1 entries
        QDT_HIP(hipHostMalloc(&h, kHostSetArrays * (k + 1) * sizeof(uint32_t), hipHostMallocDefault));
        b.host_sets = (uint32_t*)h;
        b.host_sets_cap = k + 1;
    }
    if (ws_sets_ && k <= b.sets_cap && tris <= b.tris_cap) return Status::ok;
    free_region(ws_sets_, ws_sets_bytes_);
    const size_t K = std::max(k, b.sets_cap) + 1;
    const size_t T = std::max(tris, b.tris_cap);
    const Item items[] = {
        { (void**)&b.soff, 4 * K },   { (void**)&b.srank, 4 * K },  { (void**)&b.rpos, 4 * K },
        { (void**)&b.vstart, 4 * K }, { (void**)&b.setbad, 4 * K }, { (void**)&b.in_off, 4 * K },
        { (void**)&b.mstart, 4 * K }, { (void**)&b.sstart, 4 * K }, { (void**)&b.toff, 4 * K },
        { (void**)&b.trank, 4 * T },  { (void**)&b.btris, 12 * T }, { (void**)&b.tmain, 64 },
    };
    const Status st = carve_region(items, ws_sets_, ws_sets_bytes_);
    if (st != Status::ok) return st;
    b.sets_cap = K - 1;
    b.tris_cap = T;
    return Status::ok;
}

Status Engine::run(const double* d_pts, size_t n)
{
    res_ = DeviceResult();
    batch_ = BatchResult();
    st_ = RunStats();
    if (n == 0) return Status::ok;
    Status s = ensure(n);
    if (s != Status::ok) return s;
    if (n <= kSmallMax && cells_ && !profiling_)
    {
        s = run_small(d_pts, n);
        if (s != Status::ok) res_ = DeviceResult();
        return s;
    }
    bool redo = false;
    split_ = true;
    s = run_impl(d_pts, n, nullptr, 1, seams_, redo);
    if (s == Status::ok && redo && !split_)
    {
        // a split sort met a long run out of order: repeat with full sorts
        const uint32_t bits = st_.err_bits;
        st_ = RunStats();
        s = run_impl(d_pts, n, nullptr, 1, seams_, redo);
        if (!redo) st_.err_bits = bits;
    }
    if (s == Status::ok && redo)
    {
        // a seam piece failed (budget or record exhaustion): repeat unsplit
        const RunStats first = st_;
        st_ = RunStats();
        s = run_impl(d_pts, n, nullptr, 1, false, redo);
        st_.seam_retries = 1;
        st_.device_ms += first.device_ms;
        for (int i = 0; i < 8; ++i) st_.seam[i] = first.seam[i];
        st_.err_bits = first.err_bits;
        st_.spare_records = first.spare_records;
    }
    if (s != Status::ok) res_ = DeviceResult();
    return s;
}

// A single set small enough to be one cell: k_small, one launch and one
// read of its three result words.
Status Engine::run_small(const double* d_pts, size_t n)
{
    Bufs& b = *b_;
    hipStream_t s = stream_;
    Plan P;
    P.nodes = 1;
    P.offs = 2;
    P.qv_nodes = 1;
    QDT_TRY(ensure_nodes(P));
    uint32_t* hs = b.host_scalars;
    if (ev0_) QDT_HIP(hipEventRecord(ev0_, s));
    const uint32_t nn = (uint32_t)n;
    const uint32_t tri_cap = (uint32_t)(2 * cap_n_);
    uint32_t* out = b.rblock;
    constexpr int kPer = (int)kSmallMax / 64;
    k_small<kPer><<<1, 64, std::max(cell_lds_bytes((int)kSmallMax), (size_t)24 * kSmallMax), s>>>(
        (const double2*)d_pts, nn, b.NX, b.OR, b.P0, b.VE, b.B0, reinterpret_cast<uint32_t*>(b.kA), b.ID0, b.rep,
        b.off, b.res, b.tris, tri_cap, out, b.err, b.counters);
    QDT_HIP(hipGetLastError());
    if (ev1_) QDT_HIP(hipEventRecord(ev1_, s));
    QDT_HIP(hipMemcpyAsync(hs, out, 12, hipMemcpyDeviceToHost, s));
    QDT_HIP(hipMemcpyAsync(hs + 3, b.err, 4, hipMemcpyDeviceToHost, s));
    QDT_HIP(hipStreamSynchronize(s));
    const uint32_t T = hs[0], m = hs[1], flags = hs[2], err = hs[3];
    if (flags & 1) return Status::not_finite;
    st_.err_bits = err;
    if (err & ~1u) return Status::internal;
    float ms = 0;
    if (ev0_) (void)hipEventElapsedTime(&ms, ev0_, ev1_);
    st_.device_ms = ms;
    st_.unique = m;
    st_.triangles = T;
    st_.levels = 0;
    st_.cells = 1;
    st_.block_level = 0;
    st_.cells_exact = (flags & 2) ? 1 : 0;
    st_.cell_points_max = (uint32_t)kSmallMax;
    st_.merge_lanes = (uint32_t)merge_lanes_;
    res_.n = n;
    res_.rep = b.rep;
    res_.unique = m;
    res_.tris = b.tris;
    res_.triangles = T;
    return Status::ok;
}

Status Engine::run_batch(const double* d_pts, const size_t* set_off, int k)
{
    const Status s = run_batch_impl(d_pts, set_off, k);
    st_.cell_points_max = (uint32_t)cell_max_;
    st_.merge_lanes = (uint32_t)merge_lanes_;
    return s;
}

Status Engine::run_batch_impl(const double* d_pts, const size_t* set_off, int k)
{
    res_ = DeviceResult();
    batch_ = BatchResult();
    st_ = RunStats();
    if (k <= 0) return Status::ok;
    const size_t n = set_off[k];
    batch_.k = k;
    batch_.status.assign(k, 0);
    batch_.unique.assign(k, 0);
    batch_.tri_start.assign(k, 0);
    batch_.tri_count.assign(k, 0);
    if (n == 0) return Status::ok;
    Status s = ensure(n);
    if (s != Status::ok) return s;
    s = ensure_sets((size_t)k, n);
    if (s != Status::ok) return s;
    if (!graphs_ || ext_ || profiling_) return run_batch_eager(d_pts, set_off, k);

    // the shape: the sets' sizes and the launch parameters
    std::vector<size_t> sizes(k);
    for (int i = 0; i < k; ++i) sizes[i] = set_off[i + 1] - set_off[i];
    GraphEntry* g = nullptr;
    for (GraphEntry* e : graphs_cache_)
        if (e->ws_gen == ws_gen_ && e->seams == seams_ && e->cell_max == cell_max_ &&
            e->merge_lanes == merge_lanes_ && e->sizes == sizes)
            g = e;
    if (g)
    {
        bool ok = false;
        s = replay_graph(g, d_pts, set_off, k, ok);
        if (s != Status::ok || ok) return s;
        // a guard failed (the data asks for other launches than recorded):
        // this batch runs eagerly, the graph stays for the next one
        const RunStats keep = st_;
        st_ = RunStats();
        s = run_batch_eager(d_pts, set_off, k);
        st_.graph_fallback = 1;
        st_.err_bits |= keep.err_bits;
        return s;
    }

    // a new shape: run it in record mode, then capture the same run
    GraphEntry* e = new GraphEntry();
    e->sizes = sizes;
    e->seams = seams_;
    e->cell_max = cell_max_;
    e->merge_lanes = merge_lanes_;
    mode_ = Mode::record;
    tr_ = &e->tr;
    bool redo = false;
    s = run_impl(d_pts, n, set_off, k, seams_, redo);
    mode_ = Mode::eager;
    tr_ = nullptr;
    if (s != Status::ok || redo)
    {
        delete e;
        if (s != Status::ok) return s;
        // the seam pieces ran out: this shape is run eagerly (no graph)
        return run_batch_eager(d_pts, set_off, k);
    }
    const RunStats rec = st_;
    const BatchResult rec_batch = batch_;
    const DeviceResult rec_res = res_;
    const Status cs = capture(e, d_pts, set_off, k);
    st_ = rec;
    batch_ = rec_batch;
    res_ = rec_res;
    if (cs == Status::ok)
    {
        e->ws_gen = ws_gen_;
        if (graphs_cache_.size() >= 8)
        {
            // drop the least used entry
            size_t lo = 0;
            for (size_t i = 1; i < graphs_cache_.size(); ++i)
                if (graphs_cache_[i]->uses < graphs_cache_[lo]->uses) lo = i;
            delete graphs_cache_[lo];
            graphs_cache_.erase(graphs_cache_.begin() + (long)lo);
        }
        graphs_cache_.push_back(e);
        st_.graph_captured = 1;
    }
    else
        delete e;
    return Status::ok;
}

// Ends the capture of a replay's first graph and begins its second.
Status Engine::split_capture()
{
    if (!cap_) return Status::internal;
    hipGraph_t g = nullptr;
    if (hipStreamEndCapture(stream_, &g) != hipSuccess || !g) return Status::internal;
    cap_->graph = g;
    if (hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal) != hipSuccess) return Status::internal;
    return Status::ok;
}

// The batch pipeline of a recorded shape, captured from the engine's stream
// into a graph.  The first node copies the caller's points into the
// engine's own input buffer (its source is reset before every launch), so
// the graph does not depend on where the caller keeps them.
Status Engine::capture(GraphEntry* e, const double* d_pts, const size_t* set_off, int k)
{
    const size_t n = set_off[k];
    Bufs& b = *b_;
    if (!gin_ || gin_cap_ < n)
    {
        const Item items[] = { { (void**)&gin_, 16 * std::max<size_t>(n, 64) } };
        free_region(ws_gin_, ws_gin_bytes_);
        gin_ = nullptr;
        QDT_TRY(carve_region(items, ws_gin_, ws_gin_bytes_));
        gin_cap_ = std::max<size_t>(n, 64);
    }
    Trace& t = e->tr;
    t.stage_cap = t.upload_bytes + 4096;
    if (hipHostMalloc((void**)&t.stage, t.stage_cap, hipHostMallocDefault) != hipSuccess)
    {
        t.stage = nullptr;
        return Status::out_of_memory;
    }
    t.stage_used = 0;
    if (hipHostMalloc((void**)&e->rb, 4 * (64 + (size_t)k + 1), hipHostMallocDefault) != hipSuccess)
    {
        e->rb = nullptr;
        return Status::out_of_memory;
    }
    if (t.kg == 0) return Status::internal; // nothing to triangulate: not worth a graph
    mode_ = Mode::replay;
    tr_ = &t;
    tr_cursor_ = 0;
    cap_ = e;
    hipStream_t s = stream_;
    Status st = Status::ok;
    if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess)
        st = Status::internal;
    else
    {
        if (hipMemcpyAsync(gin_, d_pts, 16 * n, hipMemcpyDeviceToDevice, s) != hipSuccess) st = Status::internal;
        bool redo = false;
        if (st == Status::ok) st = run_impl(gin_, n, set_off, k, e->seams, redo);
        hipGraph_t graph = nullptr;
        const hipError_t ee = hipStreamEndCapture(s, &graph);
        if (ee != hipSuccess || !graph) st = Status::internal;
        e->graph2 = graph;
    }
    mode_ = Mode::eager;
    tr_ = nullptr;
    cap_ = nullptr;
    if (st != Status::ok) return st;
    if (!e->graph || !e->graph2) return Status::internal;
    if (tr_cursor_ != t.sort_mask.size()) return Status::internal; // the replay took another path
    // the input copy is the graph's only root
    size_t nroot = 0;
    if (hipGraphGetRootNodes(e->graph, nullptr, &nroot) != hipSuccess || nroot != 1) return Status::internal;
    if (hipGraphGetRootNodes(e->graph, &e->in_node, &nroot) != hipSuccess) return Status::internal;
    hipGraphNodeType type;
    if (hipGraphNodeGetType(e->in_node, &type) != hipSuccess || type != hipGraphNodeTypeMemcpy) return Status::internal;
    if (hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0) != hipSuccess)
    {
        e->exec = nullptr;
        return Status::internal;
    }
    if (hipGraphInstantiate(&e->exec2, e->graph2, nullptr, nullptr, 0) != hipSuccess)
    {
        e->exec2 = nullptr;
        return Status::internal;
    }
    e->in_src = d_pts;
    (void)b;
    return Status::ok;
}

// One launch of a captured graph on this batch's points, then the result
// block: ok = false when a guard failed or the seam pieces ran out.
Status Engine::replay_graph(GraphEntry* g, const double* d_pts, const size_t* set_off, int k, bool& ok)
{
    ok = false;
    Bufs& b = *b_;
    hipStream_t s = stream_;
    const size_t n = set_off[k];
    if (d_pts != g->in_src)
    {
        QDT_HIP(hipGraphExecMemcpyNodeSetParams1D(g->exec, g->in_node, gin_, d_pts, 16 * n,
                                                  hipMemcpyDeviceToDevice));
        g->in_src = d_pts;
    }
    uint32_t* H = g->rb;
    const size_t K1 = (size_t)k + 1;
    QDT_HIP(hipEventRecord(ev0_, s));
    QDT_HIP(hipGraphLaunch(g->exec, s));
    // the guards of the first graph decide whether the recorded launches of
    // the second apply to this batch
    QDT_HIP(hipMemcpyAsync(H, &b.counters[15], 8, hipMemcpyDeviceToHost, s));
    QDT_HIP(hipStreamSynchronize(s));
    if (H[0] | H[1]) return Status::ok; // ok = false: run eagerly
    QDT_HIP(hipGraphLaunch(g->exec2, s));
    QDT_HIP(hipEventRecord(ev1_, s));
    QDT_HIP(hipMemcpyAsync(H, b.rblock, 256, hipMemcpyDeviceToHost, s));
    QDT_HIP(hipMemcpyAsync(H + 64, b.toff, 4 * K1, hipMemcpyDeviceToHost, s));
    QDT_HIP(hipStreamSynchronize(s));
    ++g->uses;
    unsigned long long cnts[16];
    std::memcpy(cnts, H + 4, sizeof cnts);
    const uint32_t T = H[0], err = H[1];
    st_.err_bits = err;
    if (cnts[15] || (err & kSeamErr)) return Status::ok; // guard failed, or the pieces ran out
    if (err & ~1u) return Status::internal;
    float ms = 0;
    (void)hipEventElapsedTime(&ms, ev0_, ev1_);
    st_.device_ms = ms;
    st_.orient = cnts[0];
    st_.orient_exact = cnts[1];
    st_.incircle = cnts[2];
    st_.incircle_exact = cnts[3];
    st_.extract_queue = cnts[4];
    for (int i = 0; i < 8; ++i) st_.seam[i] = cnts[5 + i];
    st_.merge_edges = cnts[13];
    st_.spare_records = H[2];
    st_.cells_exact = H[3];
    st_.triangles = T;
    st_.unique = g->tr.m;
    st_.graph_replay = 1;
    st_.cell_points_max = (uint32_t)cell_max_;
    st_.merge_lanes = (uint32_t)merge_lanes_;
    for (int i = 0; i < k; ++i) batch_.status[i] = g->tr.setbad[i] ? (int)Status::not_finite : 0;
    res_.n = n;
    res_.rep = b.rep;
    res_.unique = g->tr.m;
    res_.tris = b.tris;
    res_.triangles = T;
    QDT_TRY(finish_batch(H + 64, T, k, g->tr.kg, g->tr.set_rank, g->tr.vstart));
    ok = true;
    return Status::ok;
}

Status Engine::run_batch_eager(const double* d_pts, const size_t* set_off, int k)
{
    const size_t n = set_off[k];
    bool redo = false;
    Status s = run_impl(d_pts, n, set_off, k, seams_, redo);
    if (s == Status::ok && redo)
    {
        const RunStats first = st_;
        st_ = RunStats();
        s = run_impl(d_pts, n, set_off, k, false, redo);
        st_.seam_retries = 1;
        st_.device_ms += first.device_ms;
        for (int i = 0; i < 8; ++i) st_.seam[i] = first.seam[i];
        st_.err_bits = first.err_bits;
        st_.spare_records = first.spare_records;
    }
    st_.cell_points_max = (uint32_t)cell_max_;
    st_.merge_lanes = (uint32_t)merge_lanes_;
    if (s != Status::ok) batch_ = BatchResult();
    return s;
}

Status Engine::run_impl(const double* d_pts_raw, size_t n_sz, const size_t* set_off, int k, bool seams,
                        bool& redo)
{
    redo = false;
    Bufs& b = *b_;
    hipStream_t s = stream_;
    const uint32_t n = (uint32_t)n_sz;
    const double2* pts = (const double2*)d_pts_raw;
    const bool multi = set_off != nullptr;
    uint32_t* hs = b.host_scalars;
    const bool replay = mode_ == Mode::replay;
    const bool record = mode_ == Mode::record;
    unsigned long long* guard = &b.counters[15];

    // profiling: event 0..8 = stage boundaries, 9.. = merge levels (start, end)
    const bool prof = profiling_ && !replay;
    auto mark = [&](size_t i) {
        if (prof)
        {
            hipEvent_t e = prof_event(i);
            if (e) (void)hipEventRecord(e, s);
        }
    };
    if (ev0_ && !replay) QDT_HIP(hipEventRecord(ev0_, s));
    if (ext_)
    {
        // the node and seam regions are carved again after the points' region
        free_region(ws_nodes_, ws_nodes_bytes_);
        free_region(ws_anchors_, ws_anchors_bytes_);
        b.nodes_cap = b.offs_cap = b.seg_nodes = b.anchor_cap = b.qv_nodes = 0;
        ext_used_ = ext_ws_end_;
    }
    mark(0);
    fill_words(b.counters, 2 * 16, 0, s);
    fill_words(b.err, 16, 0, s);

    // 0. several sets: reject the sets holding a non-finite point, then give
    // every set its rank in the kd forest (root level ascending; roots only
    // on even levels so that every level has one cut frame)
    std::vector<int> set_rank, rank_set, roots;
    std::vector<uint32_t> rank_n;
    int kg = 1; // sets in the forest
    if (multi)
    {
        uint32_t* H = b.host_sets;
        const size_t K1 = (size_t)k + 1;
        for (int i = 0; i <= k; ++i) H[i] = (uint32_t)set_off[i];
        QDT_TRY(upload(b.soff, H, 4 * K1));
        fill_words(b.setbad, K1, 0, s);
        k_check_sets<<<cdiv(n, kBlock), kBlock, 0, s>>>(pts, n, b.soff, k, b.setbad);
        if (replay)
        {
            // the recorded verdicts, guarded against this batch's
            for (int i = 0; i < k; ++i) H[K1 + i] = tr_->setbad[i];
            QDT_TRY(upload(b.mstart, H + K1, 4 * (size_t)k));
            k_guard_words<<<cdiv((size_t)k, kBlock), kBlock, 0, s>>>(b.setbad, b.mstart, (uint32_t)k, guard, 2);
        }
        else
        {
            QDT_HIP(hipMemcpyAsync(H + K1, b.setbad, 4 * (size_t)k, hipMemcpyDeviceToHost, s));
            QDT_HIP(hipStreamSynchronize(s));
            if (record)
            {
                tr_->setbad.assign(H + K1, H + K1 + k);
                tr_->upload_bytes += 4 * (size_t)k + 512;
            }
        }
        std::vector<int> depth(k, 0);
        int D0 = 0;
        for (int i = 0; i < k; ++i)
        {
            const uint32_t ni = (uint32_t)(set_off[i + 1] - set_off[i]);
            batch_.status[i] = H[K1 + i] ? (int)Status::not_finite : 0;
            if (ni && !H[K1 + i])
            {
                depth[i] = kd_depth(ni);
                D0 = std::max(D0, depth[i]);
            }
        }
        std::vector<std::pair<int, int>> order; // (root, set)
        for (int i = 0; i < k; ++i)
        {
            const uint32_t ni = (uint32_t)(set_off[i + 1] - set_off[i]);
            if (!ni || H[K1 + i]) continue;
            int r = D0 - depth[i];
            if (r & 1) --r;
            order.push_back({ r, i });
        }
        std::sort(order.begin(), order.end());
        kg = (int)order.size();
        set_rank.assign(k, 0);
        for (int r = 0; r < kg; ++r)
        {
            set_rank[order[r].second] = r;
            rank_set.push_back(order[r].second);
            roots.push_back(order[r].first);
        }
        for (int i = 0, r = kg; i < k; ++i) // rejected and empty sets last, never rooted
        {
            const uint32_t ni = (uint32_t)(set_off[i + 1] - set_off[i]);
            if (!ni || H[K1 + i])
            {
                set_rank[i] = r++;
                rank_set.push_back(i);
            }
        }
        if (record)
        {
            tr_->set_rank = set_rank;
            tr_->kg = kg;
        }
        if (kg == 0) return Status::ok;
        uint32_t* Hr = H + K1;       // srank
        uint32_t* Hp = H + 2 * K1;   // rpos
        uint32_t* Hi = H + 3 * K1;   // in_off
        uint32_t pos = 0;
        for (int r = 0; r < k; ++r)
        {
            const int i = rank_set[r];
            Hr[i] = (uint32_t)r;
            Hp[r] = pos;
            Hi[r] = (uint32_t)set_off[i];
            pos += (uint32_t)(set_off[i + 1] - set_off[i]);
        }
        Hp[k] = pos;
        Hi[k] = n;
        QDT_TRY(upload(b.srank, Hr, 4 * K1));
        QDT_TRY(upload(b.rpos, Hp, 4 * K1));
        QDT_TRY(upload(b.in_off, Hi, 4 * K1));
        k_input_rank<<<cdiv(n, kBlock), kBlock, 0, s>>>(n, b.soff, k, b.srank, b.rin);
    }

    // 1. keys; the first sort orders by y and carries the input index
    k_keys<<<cdiv(n, kBlock), kBlock, 0, s>>>(pts, n, b.kx, b.kA, b.vA, b.err);
    uint64_t *kA = b.kA, *kB = b.kB;
    uint32_t *vA = b.vA, *vB = b.vB;
    // One set run eagerly: the digit histograms of both sorts come first (the
    // second sort's keys are the x keys in another order, so their histogram
    // is that of the x keys), and one read-back gives both sorts' passes and
    // the non-finite flag, reported before any further work.
    const bool one_read = !multi && mode_ == Mode::eager && n > 1;
    const bool split = one_read && split_ && n >= kSplitMin;
    // the radix passes of mask (ghist: the keys' digit histograms), split
    // into the top passes and a fix-up where that applies
    auto split_passes = [&](uint64_t*& ka, uint32_t*& va, uint64_t*& kb, uint32_t*& vb, uint32_t mask,
                            const uint32_t* gh) {
        uint32_t top = mask;
        if (split)
            while (__builtin_popcount(top) > kSplitPasses) top &= top - 1; // the lowest passes go
        sort_passes(ka, va, kb, vb, n, top, gh, b.hist, (uint32_t*)b.err, kSortErr, s);
        if (top != mask)
            k_sort_fixup<<<cdiv(n, kBlock), kBlock, 0, s>>>(ka, va, n, 8 * __builtin_ctz(top), (uint32_t*)b.err,
                                                             kSplitErr);
    };
    uint32_t mask_x = 0;
    Status st = Status::ok;
    if (one_read)
    {
        fill_words(b.ghist, 2 * 8 * 256, 0, s);
        const uint32_t hb = std::min<uint32_t>(cdiv(n, kBlock), 1024);
        k_radix_global_hist<<<hb, kBlock, 0, s>>>(kA, n, b.ghist);
        k_radix_global_hist<<<hb, kBlock, 0, s>>>(b.kx, n, b.ghist + 8 * 256);
        QDT_HIP(hipMemcpyAsync(b.host_hist, b.ghist, 4 * 2 * 8 * 256, hipMemcpyDeviceToHost, s));
        QDT_HIP(hipMemcpyAsync(hs, b.err, 4, hipMemcpyDeviceToHost, s));
        QDT_HIP(hipStreamSynchronize(s));
        if (hs[0] & 1) return Status::not_finite;
        mask_x = pass_mask(b.host_hist + 8 * 256, n);
        split_passes(kA, vA, kB, vB, pass_mask(b.host_hist, n), b.ghist);
    }
    else
        st = sort_keys(kA, vA, kB, vB, n);
    if (st != Status::ok) return st;
    // keep the y order for frame 1, runs of equal y put in x-descending
    // order (counters[14]: a run too long for that)
    k_ytie<<<cdiv(n, kBlock), kBlock, 0, s>>>(vA, n, b.Bt);
    k_yfix<<<cdiv(n, kBlock), kBlock, 0, s>>>(kA, b.kx, n, b.Bt, &b.counters[14]);
    if (!multi && !one_read)
    {
        // non-finite input is reported before any further work
        QDT_HIP(hipMemcpyAsync(hs, b.err, 4, hipMemcpyDeviceToHost, s));
        QDT_HIP(hipStreamSynchronize(s));
        if (hs[0] & 1) return Status::not_finite;
    }
    // (the first sort's look-back reports into err: kept until the run's end)
    if (!one_read) fill_words(b.err, 16, 0, s);
    mark(1);
    // 2. stable sort by x -> (x, y, index); several sets: then by rank
    k_gather_u64<<<cdiv(n, kBlock), kBlock, 0, s>>>(vA, b.kx, n, kA);
    if (one_read)
        split_passes(kA, vA, kB, vB, mask_x, b.ghist + 8 * 256);
    else
        st = sort_keys(kA, vA, kB, vB, n);
    if (st != Status::ok) return st;
    if (multi)
    {
        k_gather_rank<<<cdiv(n, kBlock), kBlock, 0, s>>>(vA, b.rin, n, kA);
        st = sort_keys(kA, vA, kB, vB, n);
        if (st != Status::ok) return st;
    }

    mark(2);
    // 3. dedupe
    // (several sets: the last sort's keys are the set ranks)
    const uint64_t* srank = multi ? kA : nullptr;
    k_dedupe_flags<<<cdiv(n, kBlock), kBlock, 0, s>>>(pts, vA, srank, n, b.flagp, b.P1);
    scan_u32(LoadU8{ b.flagp }, n, StoreU32{ b.dist }, b.partials, s);
    const uint32_t nbs = cdiv(n, kScanTile);
    k_dedupe_write<<<cdiv(n, kBlock), kBlock, 0, s>>>(b.P1, vA, b.flagp, srank, b.dist, n, b.P0, b.ID0, b.vrank,
                                                      b.Z);
    k_rep<<<cdiv(n, kBlock), kBlock, 0, s>>>(vA, b.flagp, b.dist, n, b.ID0, b.partials + nbs, b.rep);
    std::vector<uint32_t> vstart_h;
    if (multi)
        k_vstart<<<cdiv((size_t)k + 1, kBlock), kBlock, 0, s>>>(b.dist, b.rpos, k, n, b.partials + nbs, b.vstart);
    unsigned long long ytie = 1;
    uint32_t m;
    if (replay)
    {
        // the recorded distinct counts and frame-1 method, guarded
        m = tr_->m;
        ytie = tr_->ytie ? 1 : 0;
        vstart_h = tr_->vstart;
        k_guard_scalar<<<1, 64, 0, s>>>(b.partials + nbs, m, guard, 4);
        k_guard_flag<<<1, 64, 0, s>>>(&b.counters[14], tr_->ytie ? 1 : 0, guard, 8);
        if (multi)
        {
            QDT_TRY(upload(b.sstart, vstart_h.data(), 4 * ((size_t)k + 1)));
            k_guard_words<<<cdiv((size_t)k + 1, kBlock), kBlock, 0, s>>>(b.vstart, b.sstart, (uint32_t)k + 1,
                                                                         guard, 16);
        }
        // everything after this is sized by the recorded counts: it goes
        // into the second graph, launched only when these guards hold
        QDT_TRY(split_capture());
    }
    else
    {
        QDT_HIP(hipMemcpyAsync(hs, b.partials + nbs, 4, hipMemcpyDeviceToHost, s));
        if (multi) QDT_HIP(hipMemcpyAsync(b.host_sets, b.vstart, 4 * ((size_t)k + 1), hipMemcpyDeviceToHost, s));
        QDT_HIP(hipMemcpyAsync(hs + 4, &b.counters[14], 8, hipMemcpyDeviceToHost, s));
        // split sorts: a fix-up that met a long run out of order is known
        // here, before the tree is built on the wrong order
        if (split) QDT_HIP(hipMemcpyAsync(hs + 2, b.err, 4, hipMemcpyDeviceToHost, s));
        QDT_HIP(hipStreamSynchronize(s));
        if (split && (hs[2] & kSplitErr))
        {
            st_.err_bits = hs[2];
            split_ = false;
            redo = true;
            return Status::ok;
        }
        m = hs[0];
        memcpy(&ytie, hs + 4, 8);
        if (multi) vstart_h.assign(b.host_sets, b.host_sets + k + 1);
        if (record)
        {
            tr_->m = m;
            tr_->ytie = ytie != 0;
            tr_->vstart = vstart_h;
            tr_->upload_bytes += 4 * ((size_t)k + 1) + 512;
        }
    }
    res_.n = n;
    res_.rep = b.rep;
    res_.unique = m;
    st_.unique = m;
    const uint32_t m_forest = multi ? vstart_h[kg] : m; // distinct points in rooted sets
    if (m_forest < 3 && !multi)
    {
        float ms = 0;
        if (ev1_)
        {
            QDT_HIP(hipEventRecord(ev1_, s));
            QDT_HIP(hipEventSynchronize(ev1_));
            (void)hipEventElapsedTime(&ms, ev0_, ev1_);
        }
        else
            QDT_HIP(hipStreamSynchronize(s));
        st_.device_ms = ms;
        return Status::ok;
    }

    // the kd forest
    Plan P;
    {
        std::vector<uint32_t> sizes;
        int D;
        if (multi)
        {
            for (int r = 0; r < kg; ++r) sizes.push_back(vstart_h[r + 1] - vstart_h[r]);
            D = 0;
            for (int r = 0; r < kg; ++r) D = std::max(D, roots[r] + kd_depth(std::max<uint32_t>(sizes[r], 1)));
        }
        else
        {
            sizes.push_back(m);
            roots.assign(1, 0);
            D = kd_depth(m);
        }
        if (D + 2 > kMaxLevels) return Status::internal;
        make_plan(P, sizes, roots, D, seams);
    }
    st = ensure_nodes(P);
    if (st != Status::ok) return st;
    {
        int* T = b.host_table;
        for (int d = 0; d < kMaxLevels; ++d)
        {
            const bool in = d <= P.D;
            T[4 * d] = in ? P.count[d] : 0;
            T[4 * d + 1] = in ? P.offb[d] : 0;
            T[4 * d + 2] = in ? P.nodeb[d] : 0;
            T[4 * d + 3] = 0;
        }
        QDT_TRY(upload(b.off, T, 4 * kLvlHdr));
    }

    mark(3);
    fill_words(b.VE, m, 0xffffffffu, s);
    fill_words(b.owner, m, 0xffffffffu, s);
    // the face flags of the extraction double as the deletion marks of the
    // seam anchors (-1 = kept) until the merges end
    fill_words(b.face, b.rec_cap, 0xffffffffu, s);
    fill_words(b.spare, 16, 0, s);
    if (m_forest < m) // records of rejected sets' points stay free
        fill_words(b.OR + 6 * (size_t)m_forest, 6 * (size_t)(m - m_forest), 0xffffffffu, s);
    const int lim_e = (int)(4 * b.rec_cap);
    const int spare_base = (int)(3 * (size_t)m);
    const int spare_cap = (int)(spare_limit_ ? std::min<size_t>(spare_limit_, b.rec_cap - 3 * (size_t)m)
                                             : b.rec_cap - 3 * (size_t)m);
    // 4. frame-1 order: (y asc, x desc).  No run of equal y longer than
    // kYRunMax: the representatives of the first sort's y order with its runs
    // re-sorted (several sets: that order grouped by set, stably).
    // Otherwise a stable y-sort of the reversed frame-0 order (then by set).
    kA = b.kA;
    kB = b.kB;
    vA = b.vA;
    vB = b.vB;
    if (!ytie)
    {
        // kA, kB, vA and vB are free here (they serve the sorts of the other branch)
        const uint32_t* perm = b.Bt;
        uint32_t* w = b.vB;
        uint32_t* g = reinterpret_cast<uint32_t*>(b.kB);
        uint32_t* out = b.vA;
        if (multi)
        {
            k_rank_pairs<<<cdiv(n, kBlock), kBlock, 0, s>>>(b.Bt, b.rin, n, kA, vA);
            st = sort_keys(kA, vA, kB, vB, n, passes_upto((uint64_t)k));
            if (st != Status::ok) return st;
            perm = vA;
            w = reinterpret_cast<uint32_t*>(kA);
            g = reinterpret_cast<uint32_t*>(kB);
            out = vB;
        }
        if (m == n)
        {
            // every point distinct: each is its own representative and the
            // compaction keeps everything
            k_rep_gather<<<cdiv(n, kBlock), kBlock, 0, s>>>(perm, b.Z, n, out);
        }
        else
        {
            k_rep_id<<<cdiv(n, kBlock), kBlock, 0, s>>>(b.rep, b.Z, (uint32_t)n, w);
            k_rep_gather<<<cdiv(n, kBlock), kBlock, 0, s>>>(perm, w, n, g);
            scan_u32(RepLoad{ g }, n, RepWrite{ g, out }, b.partials, s);
        }
        vA = out;
    }
    else
    {
        k_frame1_keys<<<cdiv(m, kBlock), kBlock, 0, s>>>(b.P0, m, b.kA, b.vA);
        st = sort_keys(kA, vA, kB, vB, m, 0xffu);
        if (st != Status::ok) return st;
        if (multi)
        {
            k_gather_rank<<<cdiv(m, kBlock), kBlock, 0, s>>>(vA, b.vrank, m, kA);
            st = sort_keys(kA, vA, kB, vB, m, passes_upto((uint64_t)k));
            if (st != Status::ok) return st;
        }
    }
    uint32_t* B1 = vA;
    uint32_t* B0 = b.B0;
    // frame-1 ranks carried by the two kd arrays (and the partition's
    // output), in the sort keys' buffers
    uint32_t* R0 = reinterpret_cast<uint32_t*>(b.kA);
    uint32_t* R1 = reinterpret_cast<uint32_t*>(b.kB);
    uint32_t* Rt = reinterpret_cast<uint32_t*>(b.ky);
    // the per-node keys of each cut: in the sort's value buffer that does
    // not hold the frame-1 order (the kd arrays rotate among B0, B1, Bt)
    uint32_t* const medk = B1 == b.vA ? b.vB : b.vA;
    k_kd_init<<<cdiv(m, kBlock), kBlock, 0, s>>>(B0, B1, m, R0, R1);

    mark(4);
    // 5. kd levels
    const int D = P.D;
    // (a replay builds the recorded forest from the recorded counts: its
    // shape stays consistent with the launches even when a guard fails)
    const uint32_t* vstart = multi ? (replay ? b.sstart : b.vstart) : nullptr;
    if (!multi)
    {
        hs[2] = 0;
        hs[3] = m;
        QDT_TRY(upload(b.one_set, hs + 2, 8));
        vstart = b.one_set; // {0, m}: the one set's distinct points
    }
    st_.levels = (uint32_t)D;
    fill_words(b.res, 4 * P.nodes, 0xffffffffu, s);
    fill_words(b.lvlK, kMaxLevels, 0, s);
    uint32_t* Bt = b.Bt;
    auto roots_at = [&](int d) {
        const int ra = P.ra[d], rb = P.ra[d + 1];
        if (ra == rb) return;
        k_roots<<<cdiv((size_t)(rb - ra) + 1, kBlock), kBlock, 0, s>>>(b.off, d, ra, rb, vstart);
    };
    // cells: the first level whose nodes all fit in LDS builds everything
    // below it in one kernel (and so do the roots of sets rooted deeper)
    // (the level from which every node, the roots below included, fits a cell)
    int dc = D + 1;
    if (cells_)
        while (dc > 0 && P.smax[dc - 1] <= cell_max_) --dc;
    const bool use_cells = cells_ && dc <= D;
    if (!use_cells) dc = D;
    const int Dk = use_cells ? dc : D; // levels partitioned by the level kernels
    // a replay cuts the levels the recorded run cut, with at most its anchors
    // per node (fewer anchors than a node asks for give the same merge)
    if (replay)
    {
        if (tr_->K.size() != (size_t)D + 1) return Status::internal;
        for (int d = 0; d <= D; ++d) P.K[d] = tr_->K[d];
    }
    roots_at(0);
    // the kd levels' look-back state: two halves, used by alternate launches
    const size_t kd_half = kd_half_words(m);
    uint32_t kd_launches = 0;
    fill_words(b.hist, 2 * kd_half, 0, s);
    for (int d = 0; d < Dk; ++d)
    {
        const uint32_t cnt = (uint32_t)P.count[d];
        uint32_t*& Bf = (d & 1) ? B1 : B0;
        uint32_t*& Bo = (d & 1) ? B0 : B1;
        uint32_t* Rf = (d & 1) ? R1 : R0;
        uint32_t*& Ro = (d & 1) ? R0 : R1;
        if (cnt)
        {
            // keys in the cut's frame: ids for frame 0, frame-1 ranks for frame 1
            const uint32_t* Kf = (d & 1) ? Rf : Bf;
            const uint32_t* Ko = (d & 1) ? Ro : Bo;
            k_children<<<cdiv(cnt, kBlock), kBlock, 0, s>>>(b.off, d);
            k_node_info<<<cdiv(cnt, kBlock), kBlock, 0, s>>>(Bf, Bo, Kf, b.off, d, P.K[d], b.P0, b.splitv, b.qv,
                                                              b.kn, b.lvlK, medk, b.ext);
            if (P.K[d])
                k_node_qv<<<cdiv((size_t)cnt * P.K[d], kBlock), kBlock, 0, s>>>(Bo, b.off, d, P.K[d], b.kn, b.qv);
            (void)Ko;
            uint32_t* const half = b.hist + (kd_launches & 1) * kd_half;
            uint32_t* const other = b.hist + ((kd_launches + 1) & 1) * kd_half;
            k_kd_level<<<cdiv(m, kKdTile), kBlock, 0, s>>>(Bo, Ro, b.off, d, medk, m, Bt, Rt, half, other, b.err);
            ++kd_launches;
            std::swap(Bo, Bt);
            std::swap(Ro, Rt);
        }
        roots_at(d + 1);
    }
    if (use_cells)
        for (int d = Dk + 1; d <= D; ++d) // offsets of the roots below the cells' level
        {
            const int ra = P.ra[d], rb = P.ra[d + 1];
            if (ra < rb)
                k_roots<<<cdiv((size_t)(rb - ra) + 1, kBlock), kBlock, 0, s>>>(b.off, d, ra, rb, vstart);
        }

    // vertex ids in the cells' order; newid lives in the dedupe's x keys
    // (free since the dedupe) until the merges end, the old coordinates and
    // original indices become the spare buffers of the next run
    const uint32_t* newid = nullptr;
    if (use_cells)
    {
        uint32_t* nid = reinterpret_cast<uint32_t*>(b.kx);
        k_renum<<<cdiv(m, kBlock), kBlock, 0, s>>>(B0, m, b.P0, b.ID0, nid, b.P1, b.ID1);
        std::swap(b.P0, b.P1);
        std::swap(b.ID0, b.ID1);
        k_renum_ids<<<1024, kBlock, 0, s>>>(nid, m, b.splitv, (size_t)P.nodes);
        k_renum_ids<<<1024, kBlock, 0, s>>>(nid, m, reinterpret_cast<int*>(b.ext), 2 * (size_t)P.nodes);
        k_renum_ids<<<1024, kBlock, 0, s>>>(nid, m, b.qv, (size_t)kSegMaxK * b.qv_nodes);
        newid = nid;
    }

    // the levels the nodes asked to cut into pieces, and their anchor counts
    {
        std::vector<int> lk(kMaxLevels, 0);
        if (!replay)
        {
            // through pinned memory (the level table's upload buffer, whose
            // copy the stream has done by now)
            QDT_HIP(hipMemcpyAsync(b.host_table, b.lvlK, 4 * kMaxLevels, hipMemcpyDeviceToHost, s));
            QDT_HIP(hipStreamSynchronize(s));
            std::memcpy(lk.data(), b.host_table, 4 * kMaxLevels);
        }
        size_t seg_nodes = 1, anchor_cap = 1;
        for (int d = 0; d < D; ++d)
        {
            if (!replay)
            {
                P.K[d] = d < Dk ? std::min(P.K[d], lk[d]) : 0;
                if (P.K[d] < kSegLevelAnchors || (P.count[d] >= kSegLevelNodes && P.smax[d] < kSegLevelBig))
                    P.K[d] = 0;
            }
            if (!P.K[d]) continue;
            seg_nodes = std::max(seg_nodes, (size_t)P.count[d]);
            anchor_cap = std::max(anchor_cap, (size_t)P.count[d] * P.K[d]);
        }
        st = ensure_anchors(seg_nodes, anchor_cap);
        if (st != Status::ok) return st;
        if (record) tr_->K = P.K;
    }

    mark(5);
    // 6. leaves, 7. merges
    if (use_cells)
    {
        uint32_t ncell = (uint32_t)P.count[dc];
        for (int d = dc + 1; d <= D; ++d) ncell += (uint32_t)(P.count[d] - 2 * P.count[d - 1]);
        st_.leaf_levels = (uint32_t)(D - dc);
        st_.block_level = (uint32_t)dc;
        st_.cells = ncell;
        st_.cell_points = m_forest;
        fill_words(b.redo, 1, 0, s);
        if (ncell)
        {
            int cap = 1;
            for (int d = dc; d <= D; ++d) cap = std::max(cap, (int)P.smax[d]);
            const size_t lds = cell_lds_bytes(cap);
            // positions per lane of the cells' scans: the smallest that holds cap
            auto cell = cap <= 128 ? k_cell<2> : (cap <= 256 ? k_cell<4> : k_cell<8>);
            auto cell_x = cap <= 128 ? k_cell_exact<2> : (cap <= 256 ? k_cell_exact<4> : k_cell_exact<8>);
            cell<<<ncell, 64, lds, s>>>(b.NX, b.OR, b.P0, b.VE, B0, R0, b.off, dc, D, cap, b.res, b.redo, b.err,
                                        b.counters);
            // one workgroup per listed cell (the rest end at once): on
            // lattices nearly every cell is listed
            cell_x<<<ncell, 64, lds, s>>>(b.NX, b.OR, b.P0, b.VE, B0, R0, b.off, dc, D, cap, b.res, b.redo,
                                          b.err, b.counters);
        }
    }
    else
    {
        const uint32_t cnt = (uint32_t)P.count[D];
        k_leaves<<<cdiv(cnt, kMergeBlock), kMergeBlock, 0, s>>>(b.NX, b.OR, b.P0, b.VE, B0, B1, b.off, D, m,
                                                                 lim_e, b.res, b.err, b.counters);
    }
    mark(6);
    uint64_t kmerge_levels = 0;
    for (int d = Dk - 1; d >= 0; --d)
    {
        const uint32_t cnt = (uint32_t)P.count[d];
        mark(9 + 2 * (size_t)d);
        const int K = P.K[d];
        if (!K) kmerge_levels |= 1ull << d;
        if (!cnt)
            ;
        else if (!K)
        {
            // the tuned merges per wavefront, fewer on levels whose merges
            // would not give every SIMD two wavefronts (10M uniform: the
            // three levels 2.35 -> 2.16 ms)
            uint32_t lanes = (uint32_t)merge_lanes_;
            while (lanes > 4 && cdiv(cnt, lanes) < kMergeWavesMin) lanes >>= 1;
            auto kern = lanes >= 64 ? k_merge<64>
                                    : (lanes >= 32 ? k_merge<32>
                                                   : (lanes >= 16 ? k_merge<16> : (lanes >= 8 ? k_merge<8> : k_merge<4>)));
            kern<<<cdiv(cnt, lanes), kMergeBlock, 0, s>>>(b.NX, b.OR, b.P0, b.VE, b.off, d, m, lim_e, b.res, b.ext,
                                                          b.err, b.counters);
        }
        else
        {
            const int total = (int)(cnt * (uint32_t)K);
            k_seg_anchor<<<cdiv(2 * (size_t)total + cnt, kMergeBlock), kMergeBlock, 0, s>>>(
                b.NX, b.OR, b.P0, b.VE, b.off, b.splitv, b.qv, B0, B1, newid, b.kn, Dk, d, K, m, lim_e, b.arec, b.acoord,
                b.akey,
                b.astate, b.res, b.ext, b.nstate, b.err, b.counters);
            int sel_threads = 64;
            while (sel_threads < K) sel_threads <<= 1;
            k_seg_select<<<cnt, sel_threads, 0, s>>>(b.NX, b.OR, b.P0, b.VE, b.off, d, K, m, lim_e, b.spare,
                                                     spare_base, spare_cap, b.res, b.nstate, b.segstart, b.arec,
                                                     b.acoord, b.akey, b.astate, b.selk, b.nsel, b.alist_n, b.err,
                                                     b.counters);
            const int acap = (int)(b.anchor_cap * (2 + 4 * kCross));
            const uint32_t ga = cdiv((size_t)total, kMergeBlock);
            k_seg_mark<<<ga, kMergeBlock, 0, s>>>(b.OR, total, b.arec, b.astate, b.spare, spare_base, spare_cap,
                                                  b.owner, b.aedge, b.alist, b.alist_n, acap, b.err, b.counters);
            if (total <= kMarkSlotsMax)
                k_seg_mark_cross_slots<<<cdiv((size_t)total * 2 * kCross, kMergeBlock), kMergeBlock, 0, s>>>(
                    b.OR, total, b.arec, b.astate, b.aedge, m, lim_e, b.owner, (int*)b.face, b.alist, b.alist_n,
                    acap, b.err);
            else
                k_seg_mark_cross<<<ga, kMergeBlock, 0, s>>>(b.OR, total, b.arec, b.astate, b.aedge, m, lim_e,
                                                            b.owner, (int*)b.face, b.alist, b.alist_n, acap, b.err);
            k_seg_rebuild<<<cdiv((size_t)total * 8, kMergeBlock), kMergeBlock, 0, s>>>(
                b.NX, b.OR, b.VE, m, lim_e, b.arec, b.aedge, b.owner, (const int*)b.face, b.alist, b.alist_n, acap,
                b.err, b.counters);
            k_seg_pools<<<ga, kMergeBlock, 0, s>>>(b.NX, b.OR, total, b.arec, b.astate, b.aedge, lim_e, (int*)b.face,
                                                   b.apool);
            k_seg_compact<<<cnt, kSegThreads, 0, s>>>(d, b.selk, b.nsel, b.astate, b.aedge, b.segstart, b.nseg,
                                                   b.off);
            {
                // threads per node: the pieces of one phase, rounded up to a
                // power of two; nodes per workgroup: to fill a wavefront
                const int hp = (K + kPhases) / kPhases;
                int tpn = 1;
                while (tpn < hp) tpn <<= 1;
                tpn = std::min(tpn, kSegPieceBlock);
                // lanes per piece: doubled while the launch stays within
                // kPieceWaves wavefronts and a node within one workgroup
                int spread = 1;
                auto waves = [&](int sp) {
                    const int np = std::max(1, 64 / (tpn * sp));
                    return (size_t)cdiv(cnt, (uint32_t)np) * (size_t)std::max(1, tpn * sp * np / 64);
                };
                while (tpn * spread * 2 <= kSegPieceBlock && waves(spread * 2) <= (size_t)kPieceWaves) spread *= 2;
                const int npb = std::max(1, 64 / (tpn * spread));
                const int threads = std::max(64, tpn * spread);
                // claim table: ~48 vertices per piece of a phase, at most 64 KB
                int entries = 256;
                while (entries < 48 * npb * hp && entries < 8192) entries <<= 1;
                k_seg_pieces<<<cdiv(cnt, (uint32_t)npb), threads, 8 * (size_t)entries, s>>>(
                    b.NX, b.OR, b.P0, b.VE, b.off, d, m, lim_e, b.spare, spare_base, spare_cap, b.segstart, b.nseg,
                    tpn, npb, spread, entries, b.segpool, b.segresume, b.err, b.counters);
            }
            k_seg_final<<<cnt, kSegThreads, 0, s>>>(b.NX, b.OR, b.P0, b.VE, b.off, d, K, m,
                                                                        lim_e, b.nseg, b.nstate, b.segpool,
                                                                        b.astate, b.apool, b.res, b.err, b.counters);
        }
        mark(10 + 2 * (size_t)d);
    }
    mark(7);

    // 8. extraction
    const uint32_t nrec = (uint32_t)(3 * (size_t)m + spare_cap);
    k_extract_test<<<cdiv(nrec, kBlock), kBlock, 0, s>>>(b.NX, b.OR, b.P0, nrec, m, lim_e, b.spare,
                                                          (uint32_t)spare_cap, b.face, b.queue, b.qcap, b.counters,
                                                          b.err);
    k_extract_exact<<<256, kBlock, 0, s>>>(b.NX, b.OR, b.P0, m, lim_e, b.queue, b.counters, b.qcap, b.face);
    k_extract_overflow<<<1024, kBlock, 0, s>>>(b.NX, b.OR, b.P0, nrec, m, lim_e, b.counters, b.qcap, b.face);
    scan_u32(FaceCount{ b.face }, nrec, FacePrefix{ b.face }, b.partials, s);
    k_face_write<<<cdiv(nrec, kBlock), kBlock, 0, s>>>(b.face, b.NX, b.OR, b.ID0, nrec, m, lim_e,
                                                        multi ? 3 * m : 0xffffffffu, multi ? b.tmain : b.partials,
                                                        b.tris, (uint32_t)(2 * cap_n_), b.err);
    mark(8);
    const uint32_t nbx = cdiv(nrec, kScanTile);
    if (replay)
    {
        // the counts stay on the device; one block of scalars is read after
        // the graph (with the ranks' triangle offsets)
        QDT_TRY(batch_output(b.partials + nbx, 0, k, kg, set_rank, vstart_h));
        k_result_block<<<1, 64, 0, s>>>(b.partials + nbx, b.err, b.spare, use_cells ? b.redo : nullptr, b.counters,
                                        b.rblock);
        return Status::ok;
    }
    if (ev1_) QDT_HIP(hipEventRecord(ev1_, s));
    // the run's scalars gathered on the device, read back in one copy
    k_result_block<<<1, 64, 0, s>>>(b.partials + nbx, b.err, b.spare, use_cells ? b.redo : nullptr, b.counters,
                                    b.rblock);
    QDT_HIP(hipMemcpyAsync(hs, b.rblock, 4 * (4 + 2 * 16), hipMemcpyDeviceToHost, s));
    QDT_HIP(hipStreamSynchronize(s));
    unsigned long long cnts[16];
    std::memcpy(cnts, hs + 4, sizeof(cnts));
    st_.spare_records = hs[2];
    st_.cells_exact = use_cells ? hs[3] : 0;
    if (hs[1] & kSplitErr)
    {
        st_.err_bits = hs[1];
        split_ = false;
        redo = true;
        return Status::ok;
    }
    if (hs[1] & kSeamErr)
    {
        for (int i = 0; i < 8; ++i) st_.seam[i] = cnts[5 + i];
        st_.err_bits = hs[1];
        redo = true;
        return Status::ok;
    }
    if (hs[1] & ~1)
    {
        st_.err_bits = hs[1];
        return Status::internal;
    }
    float ms = 0;
    if (ev0_) (void)hipEventElapsedTime(&ms, ev0_, ev1_);
    st_.device_ms = ms;
    st_.orient = cnts[0];
    st_.orient_exact = cnts[1];
    st_.incircle = cnts[2];
    st_.incircle_exact = cnts[3];
    st_.extract_queue = cnts[4];
    for (int i = 0; i < 8; ++i) st_.seam[i] = cnts[5 + i];
    st_.merge_edges = cnts[13];
    st_.err_bits = hs[1];
    st_.triangles = hs[0];
    if (prof && evs_.size() >= 9 + 2 * (size_t)Dk)
    {
        for (int i = 0; i < 8; ++i)
        {
            float t = 0;
            (void)hipEventElapsedTime(&t, evs_[i], evs_[i + 1]);
            st_.stage_ms[i] = t;
        }
        double sum = 0, seam_sum = 0;
        uint32_t launches = 0;
        for (int d = 0; d < Dk; ++d)
        {
            float t = 0;
            (void)hipEventElapsedTime(&t, evs_[9 + 2 * d], evs_[10 + 2 * d]);
            if (kmerge_levels >> d & 1)
            {
                sum += t;
                ++launches;
            }
            else
                seam_sum += t;
        }
        st_.merge_launch_ms_sum = sum;
        st_.merge_launches = launches;
        st_.seam_level_ms_sum = seam_sum;
    }
    res_.tris = b.tris;
    res_.triangles = hs[0];
    if (multi) return batch_output(b.partials + nbx, hs[0], k, kg, set_rank, vstart_h);
    return Status::ok;
}

// Groups the triangles by set with set-local indices (batch_.tris) and makes
// rep set-local.  The counts the kernels need stay on the device (dT: the
// triangle total, b.tmain: the triangles of records below the spare region),
// so that the same launches serve a captured graph; set_rank: every set's
// rank, vstart: every rank's first distinct point (host).
Status Engine::batch_output(const uint32_t* dT, uint32_t T, int k, int kg, const std::vector<int>& set_rank,
                            const std::vector<uint32_t>& vstart)
{
    Bufs& b = *b_;
    hipStream_t s = stream_;
    uint32_t* hs = b.host_scalars;
    uint32_t* H = b.host_sets;
    const size_t K1 = (size_t)k + 1;
    const bool replay = mode_ == Mode::replay;
    unsigned long long* guard = &b.counters[15];
    uint32_t T_bound, c_bound;
    if (replay)
    {
        T_bound = tr_->T_bound;
        c_bound = tr_->c_bound;
        k_guard_le<<<1, 64, 0, s>>>(dT, T_bound, guard, 32);
        k_guard_spare<<<1, 64, 0, s>>>(dT, b.tmain, c_bound, guard, 64);
    }
    else
    {
        QDT_HIP(hipMemcpyAsync(hs, b.tmain, 4, hipMemcpyDeviceToHost, s));
        QDT_HIP(hipStreamSynchronize(s));
        const uint32_t tmain = std::min(hs[0], T);
        const uint32_t c = T - tmain; // triangles of spare records
        if (T > b.tris_cap || c > cap_n_)
        {
            st_.err_bits = 1u << 20;
            return Status::internal;
        }
        // the bounds a replay of this shape allows (the recorded run uses
        // them too, so that the replay's sorts see the same keys)
        T_bound = (uint32_t)std::min<size_t>(b.tris_cap, (size_t)T + std::max<size_t>(4096, T / 64));
        c_bound = (uint32_t)std::min<size_t>(cap_n_, (size_t)c + std::max<size_t>(1024, c / 16));
        if (mode_ == Mode::record)
        {
            tr_->T = T;
            tr_->tmain = tmain;
            tr_->T_bound = T_bound;
            tr_->c_bound = c_bound;
        }
    }
    k_tri_rank<<<cdiv(std::max<uint32_t>(T_bound, 1), kBlock), kBlock, 0, s>>>(b.tris, dT, b.rin, b.trank);
    k_rank_bounds<uint32_t><<<cdiv(K1, kBlock), kBlock, 0, s>>>(b.trank, b.tmain, T_bound, kg, b.mstart);
    uint64_t *kA = b.kA, *kB = b.kB;
    uint32_t *vA = b.vA, *vB = b.vB;
    k_spare_keys<<<cdiv(std::max<uint32_t>(c_bound, 1), kBlock), kBlock, 0, s>>>(b.trank, dT, b.tmain, c_bound,
                                                                                 (uint32_t)kg, kA, vA);
    QDT_TRY(sort_keys(kA, vA, kB, vB, c_bound, passes_upto((uint64_t)kg)));
    k_rank_bounds<uint64_t><<<cdiv(K1, kBlock), kBlock, 0, s>>>(kA, nullptr, c_bound, kg, b.sstart);
    k_tri_offsets<<<1, 1024, 0, s>>>(b.mstart, b.sstart, kg, b.toff);
    k_batch_tris<<<cdiv(std::max<uint32_t>(T_bound, 1), kBlock), kBlock, 0, s>>>(
        b.tris, dT, b.tmain, b.trank, kA, vA, b.mstart, b.sstart, b.toff, b.in_off, b.btris);
    k_rep_local<<<cdiv(res_.n, kBlock), kBlock, 0, s>>>(b.rep, (uint32_t)res_.n, b.soff, k);
    if (replay) return Status::ok; // the host reads the offsets after the graph
    QDT_HIP(hipMemcpyAsync(H, b.toff, 4 * ((size_t)kg + 1), hipMemcpyDeviceToHost, s));
    QDT_HIP(hipStreamSynchronize(s));
    return finish_batch(H, T, k, kg, set_rank, vstart);
}

// Per set, in the caller's order: distinct points and triangle range, from
// the ranks' triangle offsets toff (host, kg + 1 entries).
Status Engine::finish_batch(const uint32_t* toff, uint32_t T, int k, int kg, const std::vector<int>& set_rank,
                            const std::vector<uint32_t>& vstart)
{
    Bufs& b = *b_;
    if (toff[kg] != T)
    {
        st_.err_bits = 1u << 21;
        return Status::internal;
    }
    for (int i = 0; i < k; ++i)
    {
        const int r = set_rank[i];
        if (r < kg)
        {
            batch_.unique[i] = vstart[r + 1] - vstart[r];
            batch_.tri_start[i] = toff[r];
            batch_.tri_count[i] = toff[r + 1] - toff[r];
        }
    }
    batch_.tris = b.btris;
    batch_.rep = b.rep;
    batch_.n = res_.n;
    batch_.triangles = T;
    return Status::ok;
}

// ---------------------------------------------------------------------------
// exact verifier (host side)
// ---------------------------------------------------------------------------

// ---------------------------------------------------------------------------
// Half-edge mesh of a triangle list (GuibasStolfiDwyer_Wrapper.cpp:99-202):
// half-edge 3t+i runs from corner i of triangle t to corner i+1; twins are
// found by sorting the undirected edge keys (the reference scans each
// vertex's list of out-going half-edges); local vertices are numbered in
// order of first appearance by a scan over the half-edges.
// ---------------------------------------------------------------------------

enum MeshCount
{
    kmBoundary = 0,
    kmUnmatched,
    kmBadIndex,
    kmCount
};

__global__ __launch_bounds__(kBlock) void k_mesh_keys(const uint32_t* tris, uint32_t T, uint32_t n, uint64_t* keys,
                                                      uint32_t* vals, uint32_t* first, unsigned long long* cnt)
{
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= T) return;
    const uint32_t v[3] = { tris[3 * (size_t)t], tris[3 * (size_t)t + 1], tris[3 * (size_t)t + 2] };
    const bool ok = v[0] < n && v[1] < n && v[2] < n;
    if (!ok) atomicAdd(&cnt[kmBadIndex], 1ull);
    for (int i = 0; i < 3; ++i)
    {
        const uint32_t h = 3 * t + i;
        const uint32_t a = v[i], b = v[(i + 1) % 3];
        keys[h] = ok ? (((uint64_t)min(a, b) << 32) | max(a, b)) : kNoEdge;
        vals[h] = h | (a > b ? 0x80000000u : 0u); // the half-edge and its direction
        // first appearances only decrease, so a value read as already lower
        // (possibly stale, then higher than the current one) needs no atomic
        if (ok && first[a] > h) atomicMin(&first[a], h);
    }
}

// half-edge h is its origin's first appearance
struct FirstFlag
{
    const uint32_t* tris;
    const uint32_t* first;
    uint32_t n;
    __device__ uint32_t operator()(uint32_t h) const
    {
        const uint32_t u = tris[h];
        return u < n && first[u] == h ? 1u : 0u;
    }
};

struct FirstWrite
{
    FirstFlag f;
    int2* vert;
    int* o2l;
    // h's origin when h is its first appearance, else ~0u
    __device__ uint32_t load(uint32_t h) const { return f(h) ? f.tris[h] : ~0u; }
    __device__ void store(uint32_t h, uint32_t pre, uint32_t u) const
    {
        if (u == ~0u) return;
        vert[pre] = make_int2((int)u, (int)h);
        o2l[u] = (int)pre;
    }
};

__global__ __launch_bounds__(kBlock) void k_mesh_fill(const uint32_t* tris, uint32_t E, uint32_t n, const int* o2l,
                                                      int4* he)
{
    const uint32_t h = blockIdx.x * kBlock + threadIdx.x;
    if (h >= E) return;
    const uint32_t t = h / 3, i = h - 3 * t;
    const uint32_t u = tris[h];
    he[h] = make_int4(u < n ? o2l[u] : -1, -1, (int)(3 * t + (i + 1) % 3), (int)t);
}

// per group of equal edge keys: one half-edge is a boundary one, two with
// opposite directions are twins; anything else stays unmatched
__global__ __launch_bounds__(kBlock) void k_mesh_twins(const uint64_t* keys,
                                                       const uint32_t* vals, uint32_t E, int4* he,
                                                       unsigned long long* cnt)
{
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= E) return;
    const uint64_t key = keys[p];
    if (key == kNoEdge || (p > 0 && keys[p - 1] == key)) return;
    uint32_t g = 1;
    while (p + g < E && keys[p + g] == key) ++g;
    if (g == 1)
    {
        atomicAdd(&cnt[kmBoundary], 1ull);
        return;
    }
    if (g == 2)
    {
        const uint32_t a = vals[p], b = vals[p + 1];
        if ((a ^ b) >> 31) // opposite directions
        {
            const uint32_t ha = a & 0x7fffffffu, hb = b & 0x7fffffffu;
            he[ha].y = (int)hb;
            he[hb].y = (int)ha;
            return;
        }
    }
    atomicAdd(&cnt[kmUnmatched], (unsigned long long)g);
}

// Twins only, in the triangle-major half-edge numbering (3t+i from corner
// i to corner i+1): the start of an insertion or constraint on the host.
__global__ __launch_bounds__(kBlock) void k_twin_keys(const uint32_t* tris, uint32_t T, uint64_t* keys,
                                                      uint32_t* vals, int* twin)
{
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= T) return;
    const uint32_t v[3] = { tris[3 * (size_t)t], tris[3 * (size_t)t + 1], tris[3 * (size_t)t + 2] };
    for (int i = 0; i < 3; ++i)
    {
        const uint32_t h = 3 * t + i;
        const uint32_t a = v[i], b = v[(i + 1) % 3];
        keys[h] = ((uint64_t)min(a, b) << 32) | max(a, b);
        vals[h] = h | (a > b ? 0x80000000u : 0u);
        twin[h] = -1;
    }
}

__global__ __launch_bounds__(kBlock) void k_twin_pairs(const uint64_t* keys, const uint32_t* vals, uint32_t E,
                                                       int* twin, unsigned long long* bad)
{
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= E) return;
    const uint64_t key = keys[p];
    if (p > 0 && keys[p - 1] == key) return;
    uint32_t g = 1;
    while (p + g < E && keys[p + g] == key) ++g;
    if (g == 1) return;
    if (g == 2)
    {
        const uint32_t a = vals[p], b = vals[p + 1];
        if ((a ^ b) >> 31)
        {
            const uint32_t ha = a & 0x7fffffffu, hb = b & 0x7fffffffu;
            twin[ha] = (int)hb;
            twin[hb] = (int)ha;
            return;
        }
    }
    atomicAdd(bad, (unsigned long long)g);
}

namespace
{

// device allocations of one call, freed when it returns
struct TempBuffers
{
    std::vector<void*> ptrs;
    ~TempBuffers()
    {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <class X>
    Status get(X*& p, size_t count)
    {
        void* q = nullptr;
        const hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(X));
        if (e != hipSuccess) return e == hipErrorOutOfMemory ? Status::out_of_memory : Status::internal;
        ptrs.push_back(q);
        p = (X*)q;
        return Status::ok;
    }
};

} // namespace

Status verify_triangulation(const double* pts_raw, uint32_t n, const uint32_t* tris, uint32_t T, const uint32_t* rep,
                            hipStream_t s, VerifyReport& out)
{
    out = VerifyReport();
    out.triangles = T;
    const double2* P = (const double2*)pts_raw;
    const size_t E = 3 * (size_t)T;
    if (E > 0xffffffffull) return Status::internal;
    TempBuffers tmp;
    uint64_t *kA = nullptr, *kB = nullptr;
    uint32_t *vA = nullptr, *vB = nullptr, *state = nullptr, *ghist = nullptr;
    uint8_t* used = nullptr;
    int* nextb = nullptr;
    unsigned long long* cnt = nullptr;
    Status st;
    if ((st = tmp.get(kA, E)) != Status::ok || (st = tmp.get(kB, E)) != Status::ok ||
        (st = tmp.get(vA, E)) != Status::ok || (st = tmp.get(vB, E)) != Status::ok ||
        (st = tmp.get(state, radix_state_words(E))) != Status::ok ||
        (st = tmp.get(ghist, 8 * 256)) != Status::ok || (st = tmp.get(used, n)) != Status::ok ||
        (st = tmp.get(nextb, n)) != Status::ok || (st = tmp.get(cnt, kvCount)) != Status::ok)
        return st;
    uint32_t* host_ghist = nullptr;
    if (hipHostMalloc((void**)&host_ghist, 4 * 8 * 256, hipHostMallocDefault) != hipSuccess) return Status::out_of_memory;
    struct HostFree
    {
        void* p;
        ~HostFree() { (void)hipHostFree(p); }
    } hf{ host_ghist };
    hipEvent_t e0 = nullptr, e1 = nullptr;
    QDT_HIP(hipEventCreate(&e0));
    QDT_HIP(hipEventCreate(&e1));
    struct EvFree
    {
        hipEvent_t a, b;
        ~EvFree()
        {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } ef{ e0, e1 };
    QDT_HIP(hipEventRecord(e0, s));
    QDT_HIP(hipMemsetAsync(used, 0, std::max<size_t>(n, 1), s));
    QDT_HIP(hipMemsetAsync(nextb, 0xff, 4 * std::max<size_t>(n, 1), s));
    QDT_HIP(hipMemsetAsync(cnt, 0, 8 * kvCount, s));
    QDT_HIP(hipMemsetAsync(cnt + kvSecond, 0xff, 8, s));
    if (T)
    {
        k_ver_tris<<<cdiv(T, kBlock), kBlock, 0, s>>>(P, n, tris, T, rep, kA, vA, used, cnt);
        st = sort_pairs(kA, vA, kB, vB, (uint32_t)E, state, ghist, host_ghist, s);
        if (st != Status::ok) return st;
        k_ver_edges<<<cdiv(E, kBlock), kBlock, 0, s>>>(P, tris, kA, vA, (uint32_t)E, nextb, cnt);
    }
    if (n) k_ver_vertices<<<cdiv(n, kBlock), kBlock, 0, s>>>(P, n, rep, used, nextb, cnt);
    if (n && !T)
    {
        k_ver_second<<<cdiv(n, kBlock), kBlock, 0, s>>>(P, n, rep, cnt);
        k_ver_line<<<cdiv(n, kBlock), kBlock, 0, s>>>(P, n, rep, cnt);
    }
    QDT_HIP(hipEventRecord(e1, s));
    unsigned long long c[kvCount];
    QDT_HIP(hipMemcpyAsync(c, cnt, 8 * kvCount, hipMemcpyDeviceToHost, s));
    QDT_HIP(hipStreamSynchronize(s));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    out.device_ms = ms;
    out.boundary_edges = c[kvBoundary];
    out.vertices = c[kvVertices];
    out.bad_index = c[kvBadIndex];
    out.not_representative = c[kvNotRep];
    out.bad_orientation = c[kvBadOrient];
    out.bad_edges = c[kvBadEdges];
    out.non_delaunay = c[kvNonDelaunay];
    out.concave_boundary = c[kvConcave];
    out.unused_points = c[kvUnused];
    out.off_line = T ? 0 : c[kvOffLine];
    out.euler_ok = T == 0 ? true : (uint64_t)T + 2 + out.boundary_edges == 2 * out.vertices;
    // no triangles: right exactly when the distinct points are collinear
    out.ok = out.euler_ok && !out.bad_index && !out.not_representative && !out.bad_orientation && !out.bad_edges &&
        !out.non_delaunay && !out.concave_boundary && (T ? !out.unused_points : !out.off_line);
    return Status::ok;
}

Status build_mesh(const uint32_t* tris, uint32_t T, uint32_t n, int4* he, int2* vert, int* orig_to_local,
                  hipStream_t s, MeshReport& out)
{
    out = MeshReport();
    const size_t E = 3 * (size_t)T;
    if (E >= 0x80000000ull) return Status::internal; // half-edge ids and a direction bit share 32 bits
    out.halfedges = E;
    TempBuffers tmp;
    uint64_t *kA = nullptr, *kB = nullptr;
    uint32_t *vA = nullptr, *vB = nullptr, *state = nullptr, *partials = nullptr, *ghist = nullptr, *first = nullptr;
    unsigned long long* cnt = nullptr;
    Status st;
    if ((st = tmp.get(kA, E)) != Status::ok || (st = tmp.get(kB, E)) != Status::ok ||
        (st = tmp.get(vA, E)) != Status::ok || (st = tmp.get(vB, E)) != Status::ok ||
        (st = tmp.get(state, radix_state_words(E))) != Status::ok ||
        (st = tmp.get(partials, cdiv(std::max<size_t>(E, 1), kScanTile) + 2)) != Status::ok ||
        (st = tmp.get(ghist, 8 * 256)) != Status::ok || (st = tmp.get(first, n)) != Status::ok ||
        (st = tmp.get(cnt, kmCount)) != Status::ok)
        return st;
    uint32_t* host_ghist = nullptr;
    if (hipHostMalloc((void**)&host_ghist, 4 * 8 * 256, hipHostMallocDefault) != hipSuccess) return Status::out_of_memory;
    struct HostFree
    {
        void* p;
        ~HostFree() { (void)hipHostFree(p); }
    } hf{ host_ghist };
    hipEvent_t e0 = nullptr, e1 = nullptr;
    QDT_HIP(hipEventCreate(&e0));
    QDT_HIP(hipEventCreate(&e1));
    struct EvFree
    {
        hipEvent_t a, b;
        ~EvFree()
        {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } ef{ e0, e1 };
    QDT_HIP(hipEventRecord(e0, s));
    QDT_HIP(hipMemsetAsync(cnt, 0, 8 * kmCount, s));
    if (n)
    {
        QDT_HIP(hipMemsetAsync(first, 0xff, 4 * (size_t)n, s));
        QDT_HIP(hipMemsetAsync(orig_to_local, 0xff, 4 * (size_t)n, s));
    }
    uint32_t nv = 0;
    if (T)
    {
        k_mesh_keys<<<cdiv(T, kBlock), kBlock, 0, s>>>(tris, T, n, kA, vA, first, cnt);
        const FirstFlag ff{ tris, first, n };
        scan_u32(ff, (uint32_t)E, FirstWrite{ ff, vert, orig_to_local }, partials, s);
        const uint32_t nbs = cdiv((uint32_t)E, kScanTile);
        QDT_HIP(hipMemcpyAsync(&nv, partials + nbs, 4, hipMemcpyDeviceToHost, s));
        k_mesh_fill<<<cdiv(E, kBlock), kBlock, 0, s>>>(tris, (uint32_t)E, n, orig_to_local, he);
        st = sort_pairs(kA, vA, kB, vB, (uint32_t)E, state, ghist, host_ghist, s);
        if (st != Status::ok) return st;
        k_mesh_twins<<<cdiv(E, kBlock), kBlock, 0, s>>>(kA, vA, (uint32_t)E, he, cnt);
    }
    QDT_HIP(hipEventRecord(e1, s));
    unsigned long long c[kmCount];
    QDT_HIP(hipMemcpyAsync(c, cnt, 8 * kmCount, hipMemcpyDeviceToHost, s));
    QDT_HIP(hipStreamSynchronize(s));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    out.device_ms = ms;
    out.vertices = nv;
    out.boundary = c[kmBoundary];
    out.unmatched = c[kmUnmatched];
    out.bad_index = c[kmBadIndex];
    return Status::ok;
}

Status mesh_twins(const uint32_t* tris, uint32_t T, int* twin, hipStream_t s, uint64_t& unmatched)
{
    unmatched = 0;
    const size_t E = 3 * (size_t)T;
    if (E >= 0x80000000ull) return Status::internal;
    if (!T) return Status::ok;
    TempBuffers tmp;
    uint64_t *kA = nullptr, *kB = nullptr;
    uint32_t *vA = nullptr, *vB = nullptr, *state = nullptr, *ghist = nullptr;
    unsigned long long* cnt = nullptr;
    Status st;
    if ((st = tmp.get(kA, E)) != Status::ok || (st = tmp.get(kB, E)) != Status::ok ||
        (st = tmp.get(vA, E)) != Status::ok || (st = tmp.get(vB, E)) != Status::ok ||
        (st = tmp.get(state, radix_state_words(E))) != Status::ok ||
        (st = tmp.get(ghist, 8 * 256)) != Status::ok || (st = tmp.get(cnt, 1)) != Status::ok)
        return st;
    std::vector<uint32_t> host_ghist(8 * 256);
    QDT_HIP(hipMemsetAsync(cnt, 0, 8, s));
    k_twin_keys<<<cdiv(T, kBlock), kBlock, 0, s>>>(tris, T, kA, vA, twin);
    st = sort_pairs(kA, vA, kB, vB, (uint32_t)E, state, ghist, host_ghist.data(), s);
    if (st != Status::ok) return st;
    k_twin_pairs<<<cdiv(E, kBlock), kBlock, 0, s>>>(kA, vA, (uint32_t)E, twin, cnt);
    unsigned long long c = 0;
    QDT_HIP(hipMemcpyAsync(&c, cnt, 8, hipMemcpyDeviceToHost, s));
    QDT_HIP(hipStreamSynchronize(s));
    unmatched = c;
    return Status::ok;
}

} // namespace qdt
