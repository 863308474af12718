// primitives.hpp -- device-wide scan and LSD radix sort for gfx950.
//
// Both are written for 64-wide wavefronts: block-level scans combine a
// shuffle scan inside each wave with one LDS slot per wave, and the radix
// sort ranks keys inside a wave with eight ballots (one per digit bit), so a
// lane learns which lanes share its 8-bit digit without any LDS atomics.
// Sorts are stable, which the callers rely on: sorting by y and then by x
// yields the (x, y, index) order of the reference's loader.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>
#include <utility>

namespace qdt
{

constexpr int kBlock = 256;          // threads per block for streaming kernels
constexpr int kWaves = kBlock / 64;  // waves per block
constexpr int kScanItems = 8;        // items per thread in scans
constexpr int kScanTile = kBlock * kScanItems;
constexpr int kSortItems = 8;        // keys per lane in the radix sort
constexpr int kSortTile = kBlock * kSortItems;

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v)
{
    const uint32_t lane = lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1)
    {
        const uint32_t t = __shfl_up(v, d, 64);
        if (lane >= (uint32_t)d) v += t;
    }
    return v;
}

// Exclusive scan over the 256 threads of a block; returns the block total too.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t& total)
{
    __shared__ uint32_t wsum[kWaves];
    const uint32_t lane = lane_id();
    const uint32_t wave = threadIdx.x / 64;
    const uint32_t inc = wave_inclusive_scan(v);
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    uint32_t before = 0;
    total = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w)
    {
        const uint32_t s = wsum[w];
        if ((uint32_t)w < wave) before += s;
        total += s;
    }
    __syncthreads();
    return before + inc - v;
}

// ---------------------------------------------------------------------------
// Exclusive scan of n uint32 values given by a functor in(i); out(i, prefix)
// receives every exclusive prefix.  Three kernels: tile sums, a scan of the
// tile sums in one block, and the final per-tile scan.
// ---------------------------------------------------------------------------

template <class In>
__global__ __launch_bounds__(kBlock) void k_scan_reduce(In in, uint32_t n, uint32_t* partials)
{
    const uint32_t base = blockIdx.x * kScanTile;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
    {
        const uint32_t i = base + k * kBlock + threadIdx.x;
        if (i < n) s += in(i);
    }
    uint32_t total;
    block_exclusive_scan(s, total);
    if (threadIdx.x == 0) partials[blockIdx.x] = total;
}

// one block scans nb partials in place (exclusive); total -> partials[nb]
__global__ __launch_bounds__(1024) void k_scan_partials(uint32_t* partials, uint32_t nb)
{
    __shared__ uint32_t carry;
    __shared__ uint32_t wsum[16];
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t base = 0; base < nb; base += 1024)
    {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < nb ? partials[i] : 0;
        const uint32_t lane = lane_id(), wave = threadIdx.x / 64;
        const uint32_t inc = wave_inclusive_scan(v);
        if (lane == 63) wsum[wave] = inc;
        __syncthreads();
        uint32_t before = 0, total = 0;
        for (int w = 0; w < 16; ++w)
        {
            const uint32_t s = wsum[w];
            if ((uint32_t)w < wave) before += s;
            total += s;
        }
        const uint32_t c = carry;
        if (i < nb) partials[i] = c + before + inc - v;
        __syncthreads();
        if (threadIdx.x == 0) carry = c + total;
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[nb] = carry;
}

// LDS slot of tile item p: one pad word per 32 keeps the blocked reads
// (thread t reads 16 consecutive items) free of bank conflicts
__device__ __forceinline__ uint32_t scan_slot(uint32_t p) { return p + (p >> 5); }

// An output functor that splits into load(i) (loads only) and
// store(i, prefix, loaded) lets the scan issue all of a thread's loads before
// its stores.
template <class T, class = void>
struct SplitOut : std::false_type
{
};
template <class T>
struct SplitOut<T, std::void_t<decltype(std::declval<const T&>().load(0u))>> : std::true_type
{
};

template <class In, class Out>
__global__ __launch_bounds__(kBlock) void k_scan_down(In in, uint32_t n, const uint32_t* partials, Out out)
{
    // Items are read and written striped (consecutive lanes, consecutive
    // items: coalesced, and the functors' own gathers spread over the
    // wave), and scanned blocked through LDS (each thread owns kScanItems
    // consecutive items, so one thread-level prefix suffices).
    __shared__ uint32_t tile[kScanTile + kScanTile / 32];
    const uint32_t base = blockIdx.x * kScanTile;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
    {
        const uint32_t p = k * kBlock + threadIdx.x;
        const uint32_t i = base + p;
        tile[scan_slot(p)] = i < n ? in(i) : 0;
    }
    __syncthreads();
    uint32_t v[kScanItems];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
    {
        v[k] = tile[scan_slot(threadIdx.x * kScanItems + k)];
        s += v[k];
    }
    uint32_t total;
    uint32_t pre = block_exclusive_scan(s, total) + partials[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
    {
        tile[scan_slot(threadIdx.x * kScanItems + k)] = pre;
        pre += v[k];
    }
    __syncthreads();
    if constexpr (SplitOut<Out>::value)
    {
        // every item's loads first, then every item's stores: a store that
        // may alias a later item's load would otherwise hold that load back
        decltype(out.load(0u)) st[kScanItems];
#pragma unroll
        for (int k = 0; k < kScanItems; ++k)
        {
            const uint32_t i = base + k * kBlock + threadIdx.x;
            if (i < n) st[k] = out.load(i);
        }
#pragma unroll
        for (int k = 0; k < kScanItems; ++k)
        {
            const uint32_t p = k * kBlock + threadIdx.x;
            const uint32_t i = base + p;
            if (i < n) out.store(i, tile[scan_slot(p)], st[k]);
        }
    }
    else
    {
#pragma unroll
        for (int k = 0; k < kScanItems; ++k)
        {
            const uint32_t p = k * kBlock + threadIdx.x;
            const uint32_t i = base + p;
            if (i < n) out(i, tile[scan_slot(p)]);
        }
    }
}

// ---------------------------------------------------------------------------
// LSD radix sort pass over (uint64 key, uint32 value) pairs, 8-bit digits.
// ---------------------------------------------------------------------------

// per-block digit counts, digit-major: hist[d * nblocks + b]
__global__ __launch_bounds__(kBlock) void k_radix_hist(const uint64_t* keys, uint32_t n, int shift, uint32_t* hist)
{
    __shared__ uint32_t cnt[256];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kSortTile;
#pragma unroll 4
    for (int k = 0; k < kSortItems; ++k)
    {
        const uint32_t i = base + k * kBlock + threadIdx.x;
        if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & 255], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * gridDim.x + blockIdx.x] = cnt[threadIdx.x];
}

// all eight digit histograms of the whole array in one read (to skip passes
// whose digit is the same for every key)
__global__ __launch_bounds__(kBlock) void k_radix_global_hist(const uint64_t* keys, uint32_t n, uint32_t* ghist)
{
    __shared__ uint32_t cnt[8][256];
    for (int d = 0; d < 8; ++d) cnt[d][threadIdx.x] = 0;
    __syncthreads();
    const uint32_t stride = gridDim.x * kBlock;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += stride)
    {
        const uint64_t k = keys[i];
#pragma unroll
        for (int d = 0; d < 8; ++d) atomicAdd(&cnt[d][(k >> (8 * d)) & 255], 1u);
    }
    __syncthreads();
    for (int d = 0; d < 8; ++d)
    {
        const uint32_t c = cnt[d][threadIdx.x];
        if (c) atomicAdd(&ghist[d * 256 + threadIdx.x], c);
    }
}

__device__ __forceinline__ uint64_t lanemask_lt()
{
    const uint32_t lane = lane_id();
    return lane ? (~0ull >> (64 - lane)) : 0ull;
}

// Stable scatter.  Each wave owns a contiguous run of 64*kSortItems keys and
// walks it in 64-key steps; the 8 ballots give, for each lane, the mask of
// lanes holding the same digit.  offsets[d * nblocks + b] is the exclusive
// scan of the per-block histograms.  The block's keys are first placed in
// LDS in their sorted order, so that the global writes of each digit's run
// are contiguous (coalesced) instead of one scattered 8-byte store per key.
__global__ __launch_bounds__(kBlock) void k_radix_scatter(const uint64_t* kin, const uint32_t* vin, uint64_t* kout,
                                                          uint32_t* vout, uint32_t n, int shift,
                                                          const uint32_t* offsets)
{
    __shared__ uint32_t wcnt[kWaves][256];
    __shared__ uint32_t bbase[256];
    __shared__ uint64_t skey[kSortTile];
    __shared__ uint32_t sval[kSortTile];
    const uint32_t lane = lane_id();
    const uint32_t wave = threadIdx.x / 64;
    for (int w = 0; w < kWaves; ++w) wcnt[w][threadIdx.x] = 0;
    const uint32_t gbase = offsets[threadIdx.x * gridDim.x + blockIdx.x];
    __syncthreads();

    const uint32_t wbase = blockIdx.x * kSortTile + wave * 64 * kSortItems;
    uint64_t key[kSortItems];
    uint32_t val[kSortItems];
    uint32_t rank[kSortItems];
    const uint64_t lt = lanemask_lt();
#pragma unroll
    for (int k = 0; k < kSortItems; ++k)
    {
        const uint32_t i = wbase + k * 64 + lane;
        const bool valid = i < n;
        key[k] = valid ? kin[i] : 0;
        val[k] = valid ? vin[i] : 0;
        const uint32_t dg = (uint32_t)(key[k] >> shift) & 255;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b)
        {
            const uint64_t bal = __ballot((dg >> b) & 1);
            peers &= ((dg >> b) & 1) ? bal : ~bal;
        }
        const uint32_t before = wcnt[wave][dg];
        __builtin_amdgcn_wave_barrier();
        rank[k] = before + (uint32_t)__popcll(peers & lt);
        // the highest lane of each peer group advances the wave's counter
        const bool leader = valid && (peers >> lane) == 1ull;
        if (leader) wcnt[wave][dg] = before + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    {
        // per digit: exclusive prefix of the per-wave counts, then where the
        // digit's run starts in the block (a scan over the digits); bbase is
        // the global position of the block's run minus that local start
        uint32_t run = 0;
        for (int w = 0; w < kWaves; ++w)
        {
            const uint32_t c = wcnt[w][threadIdx.x];
            wcnt[w][threadIdx.x] = run;
            run += c;
        }
        uint32_t total;
        const uint32_t start = block_exclusive_scan(run, total);
#pragma unroll
        for (int w = 0; w < kWaves; ++w) wcnt[w][threadIdx.x] += start;
        bbase[threadIdx.x] = gbase - start;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kSortItems; ++k)
    {
        const uint32_t i = wbase + k * 64 + lane;
        if (i < n)
        {
            const uint32_t dg = (uint32_t)(key[k] >> shift) & 255;
            const uint32_t pos = wcnt[wave][dg] + rank[k];
            skey[pos] = key[k];
            sval[pos] = val[k];
        }
    }
    __syncthreads();
    const uint32_t b0 = blockIdx.x * kSortTile;
    const uint32_t cnt = n - b0 < (uint32_t)kSortTile ? n - b0 : (uint32_t)kSortTile;
    for (uint32_t i = threadIdx.x; i < cnt; i += kBlock)
    {
        const uint64_t k = skey[i];
        const uint32_t pos = bbase[(uint32_t)(k >> shift) & 255] + i;
        kout[pos] = k;
        vout[pos] = sval[i];
    }
}

} // namespace qdt
