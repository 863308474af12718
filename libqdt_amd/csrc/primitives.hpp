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
// LSD radix sort over (uint64 key, uint32 value) pairs, 8-bit digits, one
// kernel per digit.  The digit histograms of all eight passes come first,
// from one read of the keys; each pass then reads its tile once, ranks the
// keys inside the tile, and learns where its run of each digit starts from
// the tiles before it by a decoupled look-back: every tile publishes its
// digit counts as soon as it has them (an aggregate) and, once it knows the
// counts of all earlier tiles, their sum plus its own (an inclusive prefix).
// A tile reads the words of earlier tiles backwards, adding aggregates, until
// it reaches an inclusive prefix.  Tiles are numbered in the order their
// workgroups start (an atomic counter), so a tile only waits for tiles whose
// workgroups are already running and publish without waiting themselves.
// ---------------------------------------------------------------------------

constexpr int kSortItems = 16;                  // keys per lane in a radix pass
constexpr int kSortTile = kBlock * kSortItems;  // keys per tile
constexpr uint32_t kLbAgg = 1u << 30;           // look-back word: aggregate of one tile
constexpr uint32_t kLbInc = 2u << 30;           // look-back word: inclusive prefix
constexpr uint32_t kLbCount = (1u << 30) - 1;   // its count (sorts hold fewer than 2^30 keys)
constexpr uint32_t kLbSpins = 1u << 22;         // polls of one word before the wait gives up
constexpr size_t kSortMaxKeys = kLbCount;

// Look-back state of a sort: two halves (consecutive passes alternate), each
// a tile counter padded to 64 words and then 256 digit words per tile.
// Words: the whole state is cleared before the sort's first pass; each pass
// clears the other half for the next.
__host__ __device__ inline size_t radix_half_words(size_t n) { return 64 + 256 * ((n + kSortTile - 1) / kSortTile); }
__host__ __device__ inline size_t radix_state_words(size_t n) { return 2 * radix_half_words(n); }

// all eight digit histograms of the whole array in one read (the passes'
// digit offsets, and the passes whose digit is the same for every key).
// A digit that all active lanes of a wave share (the high bytes of keys of
// one sign and exponent) is counted by one lane: 64 LDS atomics on one
// address would run one after the other.
__global__ __launch_bounds__(kBlock) void k_radix_global_hist(const uint64_t* keys, uint32_t n, uint32_t* ghist)
{
    __shared__ uint32_t cnt[8][256];
    for (int d = 0; d < 8; ++d) cnt[d][threadIdx.x] = 0;
    __syncthreads();
    const uint32_t stride = gridDim.x * kBlock;
    for (uint32_t base = blockIdx.x * kBlock; base < n; base += stride)
    {
        const uint32_t i = base + threadIdx.x;
        const bool valid = i < n;
        const uint64_t k = valid ? keys[i] : 0;
        const uint64_t act = __ballot(valid);
        const uint32_t first = __builtin_amdgcn_readfirstlane((uint32_t)__ffsll((unsigned long long)act) - 1);
        const uint64_t k0 = ((uint64_t)__shfl(k, first, 64));
#pragma unroll
        for (int d = 0; d < 8; ++d)
        {
            const uint32_t dg = (uint32_t)(k >> (8 * d)) & 255;
            const uint32_t d0 = (uint32_t)(k0 >> (8 * d)) & 255;
            const uint64_t same = __ballot(valid && dg == d0);
            if (same == act)
            {
                if (lane_id() == first) atomicAdd(&cnt[d][d0], (uint32_t)__popcll(act));
            }
            else if (valid)
                atomicAdd(&cnt[d][dg], 1u);
        }
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

// One stable pass over digit (key >> shift) & 255, one tile of kSortTile keys
// per workgroup; launch exactly cdiv(n, kSortTile) workgroups.  state is this
// pass's half of the look-back state (cleared), next the other half (cleared
// here for the next pass).  Each wave ranks its contiguous run of
// 64*kSortItems keys in 64-key steps; 8 ballots give each lane the mask of
// lanes holding the same digit.  The tile's keys, then its values, are
// placed in LDS in sorted order, so each digit's run is written to HBM
// contiguously; a position outside [0, n) (only after a look-back that gave
// up) is not written.  A look-back
// that polls one word kLbSpins times sets *err and goes on (the sort's
// result is then wrong and the caller reports an internal error).
__global__ __launch_bounds__(kBlock) void k_radix_pass(const uint64_t* kin, const uint32_t* vin, uint64_t* kout,
                                                       uint32_t* vout, uint32_t n, int shift, const uint32_t* ghist,
                                                       uint32_t* state, uint32_t* next, uint32_t* err,
                                                       uint32_t err_bit)
{
    __shared__ uint32_t wcnt[kWaves][256];
    __shared__ uint32_t bbase[256];
    __shared__ uint32_t s_tile;
    __shared__ uint64_t skey[kSortTile];
    const uint32_t lane = lane_id();
    const uint32_t wave = threadIdx.x / 64;
    const uint32_t t = threadIdx.x;
    if (t == 0) s_tile = atomicAdd(state, 1u);
    for (int w = 0; w < kWaves; ++w) wcnt[w][t] = 0;
    // where digit t starts in the output: the exclusive prefix of the pass's
    // histogram (the scan's barriers also publish s_tile and wcnt)
    uint32_t total;
    const uint32_t gdig = block_exclusive_scan(ghist[(shift >> 3) * 256 + t], total);
    const uint32_t tile = s_tile;
    uint32_t* const st = state + 64;
    next[64 + (size_t)tile * 256 + t] = 0;
    if (tile == 0 && t == 0) next[0] = 0;

    const uint32_t wbase = tile * kSortTile + wave * 64 * kSortItems;
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
    }
#pragma unroll
    for (int k = 0; k < kSortItems; ++k)
    {
        const bool valid = wbase + k * 64 + lane < n;
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
    uint32_t run = 0;
    for (int w = 0; w < kWaves; ++w)
    {
        const uint32_t c = wcnt[w][t];
        wcnt[w][t] = run;
        run += c;
    }
    // publish the tile's count of digit t, then look back for the count of
    // digit t in all earlier tiles
    __hip_atomic_store(&st[(size_t)tile * 256 + t], (tile ? kLbAgg : kLbInc) | run, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t start = block_exclusive_scan(run, total);
    uint32_t prefix = 0;
    if (tile)
    {
        uint32_t j = tile - 1, spins = 0;
        for (;;)
        {
            const uint32_t w = __hip_atomic_load(&st[(size_t)j * 256 + t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (w & (kLbAgg | kLbInc))
            {
                prefix += w & kLbCount;
                if ((w & kLbInc) || j == 0) break;
                --j;
            }
            else if (++spins >= kLbSpins)
            {
                atomicOr(err, err_bit);
                break;
            }
            else
                __builtin_amdgcn_s_sleep(1);
        }
        __hip_atomic_store(&st[(size_t)tile * 256 + t], kLbInc | (prefix + run), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
#pragma unroll
    for (int w = 0; w < kWaves; ++w) wcnt[w][t] += start;
    bbase[t] = gdig + prefix - start;
    __syncthreads();
    // the keys, then the values, pass through one LDS area in sorted order
    // (the tile's position of item k stays in rank[k], its output position
    // of LDS slot t + j * kBlock in opos[j])
#pragma unroll
    for (int k = 0; k < kSortItems; ++k)
    {
        if (wbase + k * 64 + lane < n)
        {
            const uint32_t dg = (uint32_t)(key[k] >> shift) & 255;
            rank[k] += wcnt[wave][dg];
            skey[rank[k]] = key[k];
        }
    }
    __syncthreads();
    const uint32_t b0 = tile * kSortTile;
    const uint32_t cnt = n - b0 < (uint32_t)kSortTile ? n - b0 : (uint32_t)kSortTile;
    uint32_t opos[kSortItems];
#pragma unroll
    for (int j = 0; j < kSortItems; ++j)
    {
        const uint32_t i = t + j * kBlock;
        opos[j] = n;
        if (i < cnt)
        {
            const uint64_t k = skey[i];
            opos[j] = bbase[(uint32_t)(k >> shift) & 255] + i;
            if (opos[j] < n) kout[opos[j]] = k;
        }
    }
    __syncthreads();
    uint32_t* const sval = (uint32_t*)skey;
#pragma unroll
    for (int k = 0; k < kSortItems; ++k)
        if (wbase + k * 64 + lane < n) sval[rank[k]] = val[k];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kSortItems; ++j)
        if (opos[j] < n) vout[opos[j]] = sval[t + j * kBlock];
}

// Completes a sort whose radix passes covered only the key bits from `cut`
// up (cut a multiple of 8 in [8, 56]): keys that agree in those bits stand
// in their input order (the passes are stable).  The first thread of each
// run of such keys puts a run of at most kFixRun keys in order by insertion
// (stable: a key moves only past greater ones), the values along.  A thread
// that finds its key below the one before it in a longer run sets flag_bit
// in *flag: the caller then runs all the passes over the result, which is the full stable
// sort (keys equal in every bit are still in input order).  A run is only
// rewritten inside itself and all its keys share the bits from cut up, so
// the run boundaries every thread reads stay as they are meanwhile.
constexpr uint32_t kFixRun = 16;

__global__ __launch_bounds__(kBlock) void k_sort_fixup(uint64_t* keys, uint32_t* vals, uint32_t n, int cut,
                                                       uint32_t* flag, uint32_t flag_bit)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint64_t k = keys[i];
    const uint64_t hi = k >> cut;
    const uint64_t p = i ? keys[i - 1] : 0;
    if (i == 0 || (p >> cut) != hi)
    {
        // first key of a run: its length (up to kFixRun + 1) and order
        uint32_t len = 1;
        bool sorted = true;
        uint64_t prev = k;
        while (len <= kFixRun && i + len < n)
        {
            const uint64_t x = keys[i + len];
            if ((x >> cut) != hi) break;
            sorted = sorted && x >= prev;
            prev = x;
            ++len;
        }
        if (len <= kFixRun && !sorted)
            for (uint32_t a = i + 1; a < i + len; ++a)
            {
                const uint64_t x = keys[a];
                const uint32_t v = vals[a];
                uint32_t b = a;
                for (; b > i && keys[b - 1] > x; --b)
                {
                    keys[b] = keys[b - 1];
                    vals[b] = vals[b - 1];
                }
                keys[b] = x;
                vals[b] = v;
            }
    }
    else if (p > k)
    {
        // out of order inside a run (a run being sorted by its first thread
        // may show that for a moment; such a run is short): the run's length
        uint32_t len = 2;
        for (uint32_t a = i - 1; a > 0 && len <= kFixRun && (keys[a - 1] >> cut) == hi; --a) ++len;
        for (uint32_t a = i + 1; a < n && len <= kFixRun && (keys[a] >> cut) == hi; ++a) ++len;
        if (len > kFixRun) atomicOr(flag, flag_bit);
    }
}

} // namespace qdt
