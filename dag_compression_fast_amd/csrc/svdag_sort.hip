// svdag_sort.hip -- LSD radix sort for gfx950 (wave64), reduce-then-scan.
//
// Replaces the CUB DeviceRadixSort calls of the reference (SortKeys,
// src/DAGCompression.cu:336-357; SortPairs, :359-394).  Digits are up to 9
// bits wide.  No workgroup ever waits on another one: every pass is
//   k_sort_count    tile t (a fixed range of 4096 keys) -> per-digit counts
//   k_sort_scan     one workgroup per digit scans that digit's row of tile
//                   counts and adds the digit's global start (from the pass's
//                   digit totals, summed by the counting kernel)
//   k_sort_scatter  tile t ranks its keys, stages them in LDS in digit order
//                   and writes runs to offset[digit][t].  Stable ranking: the
//                   lanes of a wave sharing a digit OR their bits into a per-wave
//                   LDS word (rank = popcount below the lane plus a per-wave
//                   digit counter).  The first pass of a key-only sort needs no
//                   stability (its input order is arbitrary and later passes
//                   are stable) and takes its rank from one LDS atomicAdd on the
//                   tile's digit counter.
// Tiles are assigned statically (blockIdx.x + k * gridDim.x).
#include "svdag_device.hpp"
#include "svdag_internal.hpp"

#include <algorithm>
#include <mutex>
#include <type_traits>
#include <utility>
#include <vector>

namespace svdag {

SortPlan make_sort_plan(int lo, int hi, int max_bits, int ipt)
{
    SortPlan p{};
    p.ipt = ipt == kSortIptLarge ? kSortIptLarge : kSortIptSmall;
    const int key_bits = hi - lo;
    if (key_bits <= 0 || max_bits < 1 || max_bits > kMaxRadixBits) return p;
    p.passes = (key_bits + max_bits - 1) / max_bits;
    if (p.passes > kMaxPasses) {
        p.passes = 0;
        return p;
    }
    const int base = key_bits / p.passes, extra = key_bits % p.passes;
    uint32_t shift = uint32_t(lo);
    for (int i = 0; i < p.passes; ++i) {
        p.bits[i] = uint32_t(base + (i < extra ? 1 : 0));
        p.shift[i] = shift;
        shift += p.bits[i];
    }
    return p;
}

namespace {

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own L2).
// Tile order per workgroup: XCD x owns the contiguous tile range
// [x * per, (x + 1) * per), and its workgroups stride through that range, so
// tiles processed at the same time on one XCD are neighbours and their writes
// into a digit bucket land in the same L2 lines.
constexpr uint32_t kXcds = 8;

struct TileWalk {
    uint64_t tile, end, step;
};

__device__ __forceinline__ TileWalk xcd_tiles(uint64_t ntiles)
{
    const uint32_t groups = gridDim.x / kXcds;  // workgroups per XCD (grid is a multiple of kXcds)
    const uint32_t x = blockIdx.x % kXcds, local = blockIdx.x / kXcds;
    const uint64_t per = (ntiles + kXcds - 1) / kXcds;
    const uint64_t start = uint64_t(x) * per;
    const uint64_t end = min(ntiles, start + per);
    return TileWalk{start + local, end, groups};
}

// Partition of the tiles among the workgroups of a launch, each taking one
// contiguous run: workgroup b owns partition
// (b % kXcds) * (gridDim.x / kXcds) + b / kXcds when the grid is a multiple of
// kXcds (neighbouring partitions then sit on one XCD), else partition b.
// Partition q holds tiles [q * ntiles / P, (q + 1) * ntiles / P), P = gridDim.x.
struct PartRange {
    uint64_t begin, end;
};
__device__ __forceinline__ PartRange part_tiles(uint64_t ntiles)
{
    const uint64_t P = gridDim.x;
    const uint64_t q = P % kXcds ? blockIdx.x : uint64_t(blockIdx.x % kXcds) * (P / kXcds) + blockIdx.x / kXcds;
    return PartRange{q * ntiles / P, (q + 1) * ntiles / P};
}

// Chunk of input position i: the c with offsets[c] <= i < offsets[c + 1],
// searched in [lo, hi).
__device__ __forceinline__ uint64_t chunk_search(const ChunkPrefix& p, uint64_t i, uint64_t lo, uint64_t hi)
{
    while (hi - lo > 1) {
        const uint64_t mid = (lo + hi) / 2;
        if (p.offsets[mid] <= i) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Chunks of a tile's first and last key (the same for every thread).
struct TileChunks {
    uint64_t first, last;
};
__device__ __forceinline__ TileChunks tile_chunks(const ChunkPrefix& p, uint64_t base, uint32_t cnt)
{
    const uint64_t first = chunk_search(p, base, 0, p.num_chunks);
    return TileChunks{first, chunk_search(p, base + cnt - 1, first, p.num_chunks)};
}

// Global key of input key k at position i of a tile with chunks tc.
__device__ __forceinline__ uint64_t prefixed(const ChunkPrefix& p, const TileChunks& tc, uint64_t i, uint64_t k)
{
    const uint64_t c = tc.first == tc.last ? tc.first : chunk_search(p, i, tc.first, tc.last + 1);
    return (c << p.chunk_bits) | k;
}

struct HistPlan {
    int passes;
    uint32_t shift[kMaxPasses];
    uint32_t mask[kMaxPasses];
};

// Tiles whose counts the histogram writes together.
constexpr uint32_t kHistGroup = 8;

// One read of the keys gives the first pass's per-tile digit counts
// (counts[d * stride + tile]) and digit totals, and the OR of the keys (range
// check).  Each workgroup counts one contiguous run of tiles (part_tiles),
// kHistGroup tiles at a time, and writes the group's counts digit by digit as
// runs of consecutive words: one 4-byte store per tile and digit, scattered
// over the digit rows, took a third of the kernel's time.
template <typename K, int IPT>
__global__ __launch_bounds__(kSortThreads) void k_histogram(const K* __restrict__ keys,
                                                            const uint64_t* __restrict__ n_ptr, HistPlan plan,
                                                            uint32_t* __restrict__ hist,
                                                            unsigned long long* __restrict__ key_or,
                                                            uint32_t* __restrict__ counts, uint64_t stride,
                                                            ChunkPrefix prefix)
{
    constexpr uint64_t kTile = uint64_t(kSortThreads) * IPT;
    constexpr int DPT = kRadix / kSortThreads;  // digits whose totals a thread keeps
    // One row of counters per tile of the group, padded by a word so that the
    // digit-major read-out (lanes on consecutive tiles of a digit) spreads
    // over the LDS banks.
    constexpr uint32_t kRow = kRadix + 1;
    __shared__ uint32_t s_cnt[kHistGroup * kRow];
    const uint64_t n = *n_ptr;
    const PartRange part = part_tiles((n + kTile - 1) / kTile);
    const uint32_t radix0 = plan.mask[0] + 1;
    uint32_t tot[DPT] = {};
    K acc = 0;
    for (uint64_t g0 = part.begin; g0 < part.end; g0 += kHistGroup) {
        const uint32_t gn = uint32_t(min(uint64_t(kHistGroup), part.end - g0));
        for (uint32_t i = threadIdx.x; i < gn * kRow; i += kSortThreads) s_cnt[i] = 0;
        __syncthreads();
        for (uint32_t g = 0; g < gn; ++g) {
            const uint64_t base = (g0 + g) * kTile;
            const uint32_t cnt = uint32_t(min(kTile, n - base));
            // The tile's chunks are looked up before its keys are loaded,
            // whose loads would otherwise be waited for with the lookup's.
            const TileChunks tc = prefix.offsets ? tile_chunks(prefix, base, cnt) : TileChunks{0, 0};
            // Indices past the end read the last key (ignored below), so the
            // loads need no guard.  Each key is read once: non-temporal loads.
            K k[IPT];
#pragma unroll
            for (int i = 0; i < IPT; ++i)
                k[i] = __builtin_nontemporal_load(keys + min(base + uint64_t(i) * kSortThreads + threadIdx.x, n - 1));
            if (prefix.offsets) {
                // Chunk-local keys: range check, then the chunk index above them.
                bool bad = false;
#pragma unroll
                for (int i = 0; i < IPT; ++i) {
                    const uint32_t idx = uint32_t(i) * kSortThreads + threadIdx.x;
                    if (idx < cnt) {
                        bad |= (uint64_t(k[i]) >> prefix.chunk_bits) != 0;
                        k[i] = K(prefixed(prefix, tc, base + idx, uint64_t(k[i])));
                    }
                }
                if (bad) atomicOr(prefix.err, (unsigned long long)kErrKeyRange);
            }
            uint32_t* const row = s_cnt + g * kRow;
#pragma unroll
            for (int i = 0; i < IPT; ++i) {
                if (uint32_t(i) * kSortThreads + threadIdx.x < cnt) {
                    acc |= k[i];
                    atomicAdd(&row[uint32_t(k[i] >> plan.shift[0]) & plan.mask[0]], 1u);
                }
            }
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < gn * radix0; i += kSortThreads) {
            const uint32_t d = i / gn, g = i - d * gn;
            counts[uint64_t(d) * stride + g0 + g] = s_cnt[g * kRow + d];
        }
#pragma unroll
        for (int j = 0; j < DPT; ++j) {
            const uint32_t d = threadIdx.x * DPT + j;
            if (d < radix0)
                for (uint32_t g = 0; g < gn; ++g) tot[j] += s_cnt[g * kRow + d];
        }
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < DPT; ++j) {
        const uint32_t d = threadIdx.x * DPT + j;
        if (d < radix0 && tot[j]) atomicAdd(&hist[d], tot[j]);
    }
    if (key_or) {
        uint64_t a = uint64_t(acc);
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) a |= __shfl_xor(a, d, kWave);
        if (lane_id() == 0 && a) atomicOr(key_or, (unsigned long long)a);
    }
}

// Workgroups of k_histogram<K, IPT> over stride tiles: one per workgroup slot
// of the device (every partition runs at once), a multiple of kXcds, at most
// one per tile.
template <typename K, int IPT>
uint32_t histogram_grid(uint64_t stride)
{
    const uint64_t slots = uint64_t(kernel_slots(k_histogram<K, IPT>, kSortThreads));
    return uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(slots >= kXcds ? slots / kXcds * kXcds : slots, stride)));
}

// Per-tile digit counts, digit-major: counts[d * stride + tile], and the
// pass's digit totals added into hist[d].  Tiles are walked XCD by XCD, so the
// counts of neighbouring tiles are written from one L2.
template <typename K, int IPT>
__global__ __launch_bounds__(kSortThreads) void k_sort_count(const K* __restrict__ keys,
                                                             const uint64_t* __restrict__ n_ptr, uint32_t shift,
                                                             uint32_t digit_mask, uint32_t* __restrict__ counts,
                                                             uint64_t stride, uint32_t* __restrict__ hist)
{
    constexpr uint64_t kTile = uint64_t(kSortThreads) * IPT;
    __shared__ uint32_t s_cnt[kRadix];
    __shared__ uint32_t s_tot[kRadix];
    for (int i = threadIdx.x; i < kRadix; i += kSortThreads) s_tot[i] = 0;
    const uint64_t n = *n_ptr;
    const uint64_t ntiles = (n + kTile - 1) / kTile;
    const uint32_t radix = digit_mask + 1;
    const TileWalk walk = xcd_tiles(ntiles);
    // Keys of tile t; indices past the end read the last key (counted as
    // invalid below), so the loads need no guard.
    auto load = [&](uint64_t t, K (&k)[IPT]) {
        const uint64_t base = t * kTile;
#pragma unroll
        for (int i = 0; i < IPT; ++i)
            k[i] = __builtin_nontemporal_load(keys + min(base + uint64_t(i) * kSortThreads + threadIdx.x, n - 1));
    };
    K cur[IPT], nxt[IPT];
    if (walk.tile < walk.end) load(walk.tile, cur);
    // No load may be pending at the loop head (see the wait at its end).
    __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
    for (uint64_t tile = walk.tile; tile < walk.end; tile += walk.step) {
        for (uint32_t d = threadIdx.x; d < radix; d += kSortThreads) s_cnt[d] = 0;
        __syncthreads();
        // The next tile's keys are in flight while this one is counted.
        const bool more = tile + walk.step < walk.end;
        if (more) load(tile + walk.step, nxt);
        const uint64_t base = tile * kTile;
#pragma unroll
        for (int i = 0; i < IPT; ++i)
            if (base + uint64_t(i) * kSortThreads + threadIdx.x < n)
                atomicAdd(&s_cnt[uint32_t(cur[i] >> shift) & digit_mask], 1u);
        __syncthreads();
        for (uint32_t d = threadIdx.x; d < radix; d += kSortThreads) {
            const uint32_t c = s_cnt[d];
            counts[uint64_t(d) * stride + tile] = c;
            s_tot[d] += c;
        }
        __syncthreads();
        __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
        if (more) {
#pragma unroll
            for (int i = 0; i < IPT; ++i) cur[i] = nxt[i];
        }
    }
    for (uint32_t d = threadIdx.x; d < radix; d += kSortThreads) {
        const uint32_t v = s_tot[d];
        if (v) atomicAdd(&hist[d], v);
    }
}

// Block d: exclusive scan of row d of the tile counts, offset by the digit's
// global start (sum of the pass histogram below d).  In place.  Rows of at
// most kScanItems waves' worth of tiles take a one-wave block.
constexpr int kScanThreads = 1024;
constexpr int kScanItems = 4;
constexpr uint64_t kSmallScanTiles = uint64_t(kWave) * kScanItems;

template <int kScanThreads>
__global__ __launch_bounds__(kScanThreads) void k_sort_scan(uint32_t* __restrict__ counts,
                                                            const uint32_t* __restrict__ hist,
                                                            const uint64_t* __restrict__ n_ptr, uint64_t stride,
                                                            uint64_t tile_keys)
{
    __shared__ uint64_t s_red[kScanThreads / kWave];
    __shared__ uint32_t s_scan[kScanThreads / kWave];
    const uint32_t d = blockIdx.x;
    const uint64_t n = *n_ptr;
    const uint64_t ntiles = (n + tile_keys - 1) / tile_keys;
    // Global start of digit d.
    uint64_t below = 0;
    for (uint32_t i = threadIdx.x; i < d; i += kScanThreads) below += hist[i];
    uint64_t running = block_reduce_sum<kScanThreads>(below, s_red);
    uint32_t* row = counts + uint64_t(d) * stride;
    // kScanItems consecutive counts per thread per block scan.
    for (uint64_t chunk = 0; chunk < ntiles; chunk += uint64_t(kScanThreads) * kScanItems) {
        const uint64_t i0 = chunk + uint64_t(threadIdx.x) * kScanItems;
        uint32_t v[kScanItems];
        uint32_t sum = 0;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            v[k] = i0 + k < ntiles ? row[i0 + k] : 0u;
            sum += v[k];
        }
        uint32_t total;
        uint64_t run = running + block_exclusive_scan<kScanThreads>(sum, s_scan, total);
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            if (i0 + k < ntiles) row[i0 + k] = uint32_t(run);
            run += v[k];
        }
        running += total;
    }
}

// Ranking of a scatter pass: unstable (the first pass of a key-only sort),
// stable, or adaptive: unstable for a tile whose keys all agree on the bits
// the earlier passes sorted (their order within a digit is then free; the
// tiles follow the input order), stable otherwise.  A stable key-only sort
// (keys equal on the sorted bits keep their input order) ranks adaptively too
// when no key has a bit at or above the sorted bits (*key_or, OR of all keys):
// keys equal on the sorted bits are then identical; otherwise every tile ranks
// stably.
enum RankMode { kRankUnstable = 0, kRankStable = 1, kRankAdaptive = 2 };

// KO narrower than K (the top-digit pass of a 64-bit key sort): each key is
// written without its digit and the bits above it, as a key local to its digit.
template <typename K, bool PAIRS, int MODE, int IPT, int THREADS, typename KO = K>
__global__ __launch_bounds__(THREADS) void k_sort_scatter(const K* __restrict__ keys_in, KO* __restrict__ keys_out,
                                                               const uint32_t* __restrict__ vals_in,
                                                               uint32_t* __restrict__ vals_out,
                                                               const uint64_t* __restrict__ n_ptr, uint32_t shift,
                                                               uint32_t digit_mask,
                                                               const uint32_t* __restrict__ offsets, uint64_t stride,
                                                               ChunkPrefix prefix, uint32_t sorted_shift,
                                                               uint32_t sorted_mask,
                                                               const unsigned long long* __restrict__ key_or,
                                                               uint32_t end_bit)
{
    constexpr bool MATCH = MODE != kRankUnstable;  // needs the stable ranking's tables
    constexpr bool DCNT = MODE != kRankStable;     // needs the unstable ranking's counters
    constexpr int W = THREADS / kWave;
    constexpr uint32_t kTile = uint32_t(THREADS) * IPT;
    constexpr int DPT = kRadix / THREADS;  // digit slots per thread in the tile scan
    // Per-wave digit counts and tile-local digit starts never exceed the tile
    // size, so 16 bits hold them.
    static_assert(kTile <= 65535, "16-bit tile-local counters");
    __shared__ K s_keys[kTile];
    __shared__ uint32_t s_vals[PAIRS ? kTile : 1];
    // Stable ranking: per-wave digit counts and the per-wave match table (zeroed
    // once; each key's ranking clears the word it set, so it is zero again at
    // the end of every tile).  Unstable ranking: the tile's digit counts.
    __shared__ uint16_t s_wcnt[MATCH ? W : 1][kRadix];
    __shared__ uint64_t s_match_all[MATCH ? W * kRadix : 1];
    __shared__ uint32_t s_dcnt[DCNT ? kRadix : 1];
    __shared__ uint32_t s_span[2 * W];  // per wave: min and max of the sorted bits
    uint64_t* const s_match = s_match_all + (MATCH ? (threadIdx.x / kWave) * kRadix : 0);
    __shared__ uint16_t s_lstart[kRadix];
    __shared__ uint32_t s_base[kRadix];
    __shared__ uint32_t s_scan[W];

    const uint32_t tid = threadIdx.x;
    const uint32_t w = tid / kWave;
    const uint32_t lane = lane_id();
    const uint64_t n = *n_ptr;
    const uint64_t ntiles = (n + kTile - 1) / kTile;
    const uint32_t radix = digit_mask + 1;

    // Keys above the sorted bits: every tile ranks stably.
    const bool wide = MODE == kRankAdaptive && key_or && end_bit < 64 && (*key_or >> end_bit) != 0;
    const uint32_t wave_base = w * kWave * IPT;
    TileWalk walk = xcd_tiles(ntiles);
    K key[IPT];
    uint32_t val[PAIRS ? IPT : 1];
    // Keys of tile t (loaded one tile ahead).  Full tiles, the common case, load
    // without per-key guards.
    auto load_tile = [&](uint64_t t) {
        const uint64_t base = t * kTile;
        const uint32_t cnt = t < walk.end ? uint32_t(min(uint64_t(kTile), n - base)) : 0u;
        const K* src = keys_in + base + wave_base + lane;
        const uint32_t* vs = PAIRS ? vals_in + base + wave_base + lane : nullptr;
        if (cnt == kTile) {
            // Each key is read once: non-temporal loads (measured 4% faster
            // per pass than cached ones; non-temporal stores were 2x slower).
#pragma unroll
            for (int i = 0; i < IPT; ++i) {
                key[i] = __builtin_nontemporal_load(src + i * kWave);
                if constexpr (PAIRS) val[i] = vs[i * kWave];
            }
        } else {
#pragma unroll
            for (int i = 0; i < IPT; ++i) {
                const bool valid = wave_base + i * kWave + lane < cnt;
                key[i] = valid ? src[i * kWave] : K(0);
                if constexpr (PAIRS) val[i] = valid ? vs[i * kWave] : 0u;
            }
        }
        if (prefix.offsets && cnt) {
            const TileChunks tc = tile_chunks(prefix, base, cnt);
#pragma unroll
            for (int i = 0; i < IPT; ++i) {
                const uint32_t idx = wave_base + i * kWave + lane;
                if (idx < cnt) key[i] = K(prefixed(prefix, tc, base + idx, uint64_t(key[i])));
            }
        }
    };
    load_tile(walk.tile);
    if constexpr (MATCH)
        for (uint32_t i = tid; i < W * kRadix; i += THREADS) s_match_all[i] = 0;

    // Each tile: clear the counters, rank, digit starts, stage, store.  The
    // barrier after the clear also orders the previous tile's stores (reads of
    // s_keys and s_base) before this tile's writes of them.
    for (uint64_t tile = walk.tile; tile < walk.end; tile += walk.step) {
        const uint64_t tile_base = tile * kTile;
        const uint32_t tile_n = uint32_t(min(uint64_t(kTile), n - tile_base));
        if constexpr (MODE == kRankAdaptive) {
            // The span of the earlier passes' bits over the tile's keys.
            uint32_t lo = 0xFFFFFFFFu, hi = 0;
#pragma unroll
            for (int i = 0; i < IPT; ++i)
                if (wave_base + i * kWave + lane < tile_n) {
                    const uint32_t v = uint32_t(key[i] >> sorted_shift) & sorted_mask;
                    lo = min(lo, v);
                    hi = max(hi, v);
                }
#pragma unroll
            for (int o = kWave / 2; o >= 1; o >>= 1) {
                lo = min(lo, uint32_t(__shfl_xor(int(lo), o, kWave)));
                hi = max(hi, uint32_t(__shfl_xor(int(hi), o, kWave)));
            }
            if (lane == 0) {
                s_span[w] = lo;
                s_span[W + w] = hi;
            }
        }
        if constexpr (MATCH)
            for (uint32_t i = tid; i < W * kRadix; i += THREADS) (&s_wcnt[0][0])[i] = 0;
        if constexpr (DCNT)
            for (uint32_t i = tid; i < kRadix; i += THREADS) s_dcnt[i] = 0;
        __syncthreads();
        bool stable = MODE == kRankStable;
        if constexpr (MODE == kRankAdaptive) {
            uint32_t lo = s_span[0], hi = s_span[W];
#pragma unroll
            for (int i = 1; i < W; ++i) {
                lo = min(lo, s_span[i]);
                hi = max(hi, s_span[W + i]);
            }
            stable = lo != hi || wide;
        }

        // Ranking, digit starts and staging; FULL (every key of the tile present,
        // the common case) drops the per-key bounds checks and their registers.
        auto rank_and_stage = [&](auto full_tag) {
        constexpr bool FULL = decltype(full_tag)::value;
        // Per key: (rank among the wave's / tile's keys of its digit) << 16 | digit.
        uint32_t dr[IPT];
        uint16_t* wc = s_wcnt[MATCH ? w : 0];
        if (DCNT && !stable) {
#pragma unroll
            for (int i = 0; i < IPT; ++i) {
                const bool valid = FULL || wave_base + i * kWave + lane < tile_n;
                const uint32_t d = uint32_t(key[i] >> shift) & digit_mask;
                const uint32_t r = valid ? atomicAdd(&s_dcnt[d], 1u) : 0u;
                dr[i] = r << 16 | d;
            }
        }
        if (MATCH && stable) {
#pragma unroll
        for (int i = 0; i < (MATCH ? IPT : 0); ++i) {
            const bool valid = FULL || wave_base + i * kWave + lane < tile_n;
            const uint32_t d = uint32_t(key[i] >> shift) & digit_mask;
            // Lanes holding the same digit: every lane ORs its bit into the wave's
            // word for its digit (the OR does not depend on the order of the
            // lanes), reads the word back and clears it for the next key.
            uint64_t* const slot = s_match + d;
            if (valid) atomicOr((unsigned long long*)slot, 1ull << lane);
            __builtin_amdgcn_wave_barrier();
            const uint64_t match = valid ? *slot : 0ull;
            __builtin_amdgcn_wave_barrier();
            if (valid) *slot = 0ull;
            const uint32_t below = __popcll(match & lanemask_lt());
            uint32_t c = 0;
            if (valid) c = wc[d];
            // Every lane of the group reads the counter before its highest lane
            // (the leader) writes it back.
            __builtin_amdgcn_wave_barrier();
            dr[i] = (c + below) << 16 | d;
            if (valid && (match >> lane) == 1ull) wc[d] = uint16_t(c + uint32_t(__popcll(match)));
            __builtin_amdgcn_wave_barrier();
        }
        }
        __syncthreads();

        // Per digit: wave counts -> exclusive per-wave starts; tile-local digit
        // starts by a block scan over DPT consecutive digits per thread.
        uint32_t dtot[DPT];
        uint32_t tsum = 0;
#pragma unroll
        for (int k = 0; k < DPT; ++k) {
            const uint32_t d = tid * DPT + k;
            uint32_t total = 0;
            if (d < radix) {
                if (MATCH && stable) {
#pragma unroll
                    for (int i = 0; i < W; ++i) {
                        const uint32_t c = s_wcnt[i][d];
                        s_wcnt[i][d] = uint16_t(total);
                        total += c;
                    }
                } else if constexpr (DCNT) {
                    total = s_dcnt[d];
                }
            }
            dtot[k] = total;
            tsum += total;
        }
        uint32_t unused;
        uint32_t lstart = block_exclusive_scan<THREADS>(tsum, s_scan, unused);
#pragma unroll
        for (int k = 0; k < DPT; ++k) {
            const uint32_t d = tid * DPT + k;
            if (d < radix) {
                s_lstart[d] = uint16_t(lstart);
                s_base[d] = offsets[uint64_t(d) * stride + tile] - lstart;
            }
            lstart += dtot[k];
        }
        __syncthreads();

#pragma unroll
        for (int i = 0; i < IPT; ++i) {
            if (FULL || wave_base + i * kWave + lane < tile_n) {
                const uint32_t d = dr[i] & 0xFFFFu;
                const uint32_t pos = s_lstart[d] + (MATCH && stable ? uint32_t(s_wcnt[w][d]) : 0u) + (dr[i] >> 16);
                s_keys[pos] = key[i];
                if constexpr (PAIRS) s_vals[pos] = val[i];
            }
        }
        };
        if (tile_n == kTile) rank_and_stage(std::true_type{});
        else rank_and_stage(std::false_type{});
        __syncthreads();
        // Next tile's loads go out before this tile's stores.
        load_tile(tile + walk.step);
        auto store = [&](uint32_t j) {
            const K k = s_keys[j];
            const uint32_t d = uint32_t(k >> shift) & digit_mask;
            const uint32_t o = s_base[d] + j;
            if constexpr (sizeof(KO) < sizeof(K))
                keys_out[o] = KO(k & ((K(1) << shift) - 1));
            else
                keys_out[o] = k;
            if constexpr (PAIRS) vals_out[o] = s_vals[j];
        };
        if (tile_n == kTile) {
#pragma unroll
            for (int i = 0; i < IPT; ++i) store(tid + i * THREADS);
        } else {
            for (uint32_t j = tid; j < tile_n; j += THREADS) store(j);
        }
    }
}

}  // namespace

int kernel_slots(const void* kernel, int threads)
{
    static std::mutex mu;
    static std::vector<std::pair<const void*, int>> cache;
    std::lock_guard<std::mutex> lock(mu);
    for (const auto& e : cache)
        if (e.first == kernel) return e.second;
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, kernel, threads, 0) != hipSuccess || blocks < 1) blocks = 1;
    const int slots = blocks * device_cu_count();
    cache.emplace_back(kernel, slots);
    return slots;
}

int device_cu_count()
{
    static int cus = 0;
    if (cus == 0) {
        int dev = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
            cus = prop.multiProcessorCount;
        if (cus <= 0) cus = 256;
    }
    return cus;
}

constexpr unsigned kUnstableExtraLds = 20 * 1024;

template <typename K, int IPT>
static void launch_pass(hipStream_t stream, const K* src, K* dst, const uint32_t* vsrc, uint32_t* vdst,
                        const uint64_t* n_ptr, uint32_t shift, uint32_t dmask, uint32_t* counts, uint32_t* hist,
                        uint64_t stride, uint32_t grid, bool first, const ChunkPrefix& prefix, uint32_t sorted_lo,
                        bool stable, const unsigned long long* key_or, uint32_t end_bit)
{
    const bool counted = first;  // the histogram kernel counted the first pass's tiles
    // Bits the earlier passes sorted: [sorted_lo, shift).  The adaptive ranking
    // compares them as one 32-bit word, so wider spans rank stably.
    const bool span_fits = shift > sorted_lo && shift - sorted_lo <= 31;
    const uint32_t sorted_mask = span_fits ? (1u << (shift - sorted_lo)) - 1u : 0u;
    // Key sorts (tens of millions of keys) launch about one workgroup per tile
    // (up to 64 per CU): the dispatcher then balances the tiles, which measured
    // 4% faster than 8 per CU walking several tiles each.  Key/value sorts are
    // small and sized from a loose bound, where the empty workgroups would cost
    // more than they balance.
    const uint32_t scatter_grid =
        vsrc ? grid
             : uint32_t(std::max<uint64_t>(
                   kXcds, std::min<uint64_t>((stride + kXcds - 1) / kXcds * kXcds, uint64_t(device_cu_count()) * 64) /
                              kXcds * kXcds));
    if (!counted) {
        TimedLaunch t(vsrc ? kKPairCount : kKSortCount, stream);
        hipLaunchKernelGGL((k_sort_count<K, IPT>), dim3(grid), dim3(kSortThreads), 0, stream, src, n_ptr, shift, dmask,
                           counts, stride, hist);
    }
    {
        TimedLaunch t(kKSortScan, stream);
        if (stride <= kSmallScanTiles)
            hipLaunchKernelGGL(k_sort_scan<kWave>, dim3(dmask + 1), dim3(kWave), 0, stream, counts, hist, n_ptr,
                               stride, uint64_t(kSortThreads) * IPT);
        else
            hipLaunchKernelGGL(k_sort_scan<kScanThreads>, dim3(dmask + 1), dim3(kScanThreads), 0, stream, counts,
                               hist, n_ptr, stride, uint64_t(kSortThreads) * IPT);
    }
    // A stable key-only sort ranks adaptively when it can (see k_sort_scatter);
    // its first pass has no earlier bits, so every tile is uniform there.
    const unsigned long long* wide_or = stable ? key_or : nullptr;
    if (vsrc) {
        TimedLaunch t(kKSortScatterPairs, stream);
        hipLaunchKernelGGL((k_sort_scatter<K, true, kRankStable, IPT, kSortThreads>), dim3(scatter_grid),
                           dim3(kSortThreads), 0, stream, src, dst, vsrc, vdst, n_ptr, shift, dmask, counts, stride,
                           prefix, sorted_lo, sorted_mask, nullptr, 64u);
    } else if (first && !stable) {
        // Key-only first pass: the order within a digit is free.  Its kernel
        // needs fewer registers and would fit three workgroups per CU; the extra
        // LDS holds it to two, as the stable pass, which measured faster (more
        // tiles in flight spread the scattered writes over more open lines).
        TimedLaunch t(kKSortScatterKeys, stream);
        hipLaunchKernelGGL((k_sort_scatter<K, false, kRankUnstable, IPT, kSortThreads>), dim3(scatter_grid),
                           dim3(kSortThreads), kUnstableExtraLds, stream, src, dst, nullptr, nullptr, n_ptr, shift,
                           dmask, counts, stride, prefix, sorted_lo, sorted_mask, nullptr, 64u);
    } else if (span_fits || (first && wide_or)) {
        // Later key-only passes: tiles uniform in the earlier passes' bits (the
        // common case: the input is sorted on them, in runs longer than a tile)
        // rank unstably, the others stably.
        TimedLaunch t(kKSortScatterKeys, stream);
        hipLaunchKernelGGL((k_sort_scatter<K, false, kRankAdaptive, IPT, kSortThreads>), dim3(scatter_grid),
                           dim3(kSortThreads), 0, stream, src, dst, nullptr, nullptr, n_ptr, shift, dmask, counts,
                           stride, prefix, sorted_lo, sorted_mask, wide_or, end_bit);
    } else {
        TimedLaunch t(kKSortScatterKeys, stream);
        hipLaunchKernelGGL((k_sort_scatter<K, false, kRankStable, IPT, kSortThreads>), dim3(scatter_grid),
                           dim3(kSortThreads), 0, stream, src, dst, nullptr, nullptr, n_ptr, shift, dmask, counts,
                           stride, prefix, sorted_lo, sorted_mask, nullptr, 64u);
    }
}

template <typename K>
int launch_sort(hipStream_t stream, const SortPlan& plan, const K* keys_in, const uint32_t* vals_in, K* keys_a,
                K* keys_b, uint32_t* vals_a, uint32_t* vals_b, K* final_keys, uint32_t* final_vals,
                const uint64_t* n_ptr, uint64_t n_bound, uint32_t* zero, uint32_t* counts, unsigned long long* key_or,
                const ChunkPrefix& prefix)
{
    if (plan.passes == 0 || n_bound == 0) return -1;
    const bool pairs = vals_in != nullptr;
    // Key/value sorts (the unique-leaf order: a few hundred thousand keys) use
    // 2048-key tiles for u64 keys, so that more workgroups share each pass.
    const int ipt = pairs ? (sizeof(K) == 8 ? kSortIptPairs : kSortIptSmall) : plan.ipt;
    uint32_t* hist = zero;
    const uint64_t stride = sort_tiles(n_bound, ipt);
    HistPlan hp{};
    hp.passes = plan.passes;
    for (int p = 0; p < plan.passes; ++p) {
        hp.shift[p] = plan.shift[p];
        hp.mask[p] = (1u << plan.bits[p]) - 1u;
    }
    const int cus = device_cu_count();
    // A multiple of the XCD count, as the tile walk assumes.
    const uint32_t grid = uint32_t(
        std::max<uint64_t>(kXcds, std::min<uint64_t>((stride + kXcds - 1) / kXcds * kXcds, uint64_t(cus) * 8) / kXcds * kXcds));
    auto run = [&](auto ipt_tag) -> int {
        constexpr int I = decltype(ipt_tag)::value;
        {
            // Also writes the first pass's tile counts.
            TimedLaunch t(vals_in ? kKPairHistogram : kKHistogram, stream);
            hipLaunchKernelGGL((k_histogram<K, I>), dim3(histogram_grid<K, I>(stride)), dim3(kSortThreads), 0, stream,
                               keys_in, n_ptr, hp, hist, key_or, counts, stride, prefix);
        }
        const K* src = keys_in;
        const uint32_t* vsrc = vals_in;
        int where = -1;
        for (int p = 0; p < plan.passes; ++p) {
            const bool last = p + 1 == plan.passes;
            K* dst = (p % 2 == 0) ? keys_a : keys_b;
            uint32_t* vdst = (p % 2 == 0) ? vals_a : vals_b;
            where = p % 2;
            if (last && final_keys) {
                dst = final_keys;
                where = 2;
            }
            if (last && final_vals) vdst = final_vals;
            const uint32_t dmask = (1u << plan.bits[p]) - 1u;
            launch_pass<K, I>(stream, src, dst, vsrc, vdst, n_ptr, plan.shift[p], dmask, counts, hist + p * kRadix,
                              stride, grid, p == 0, p == 0 ? prefix : ChunkPrefix{}, plan.shift[0], plan.stable,
                              key_or, plan.shift[plan.passes - 1] + plan.bits[plan.passes - 1]);
            src = dst;
            vsrc = vdst;
        }
        return where;
    };
    if (ipt == kSortIptLarge) return run(std::integral_constant<int, kSortIptLarge>{});
    if constexpr (sizeof(K) == 8)
        if (ipt == kSortIptPairs) return run(std::integral_constant<int, kSortIptPairs>{});
    return run(std::integral_constant<int, kSortIptSmall>{});
}

template int launch_sort<uint64_t>(hipStream_t, const SortPlan&, const uint64_t*, const uint32_t*, uint64_t*,
                                   uint64_t*, uint32_t*, uint32_t*, uint64_t*, uint32_t*, const uint64_t*, uint64_t,
                                   uint32_t*, uint32_t*, unsigned long long*, const ChunkPrefix&);
template int launch_sort<uint32_t>(hipStream_t, const SortPlan&, const uint32_t*, const uint32_t*, uint32_t*,
                                   uint32_t*, uint32_t*, uint32_t*, uint32_t*, uint32_t*, const uint64_t*, uint64_t,
                                   uint32_t*, uint32_t*, unsigned long long*, const ChunkPrefix&);

// ---------------------------------------------------------------------------
// Segmented key sort of a batched build.  The chunks arrive one after the
// other in Morton order, so the chunk index (the top of the sorted bits) is
// already in order: sorting the chunk-local bits [lo, chunk_bits) inside each
// chunk finishes the job.  Tiles of at most kSegTile keys never cross a chunk
// (the host lays them out, chunk by chunk).  Per pass: per-tile counts of the
// pass digit (digit-major inside each chunk), a scan of each (chunk, digit)
// row over the chunk's tiles, a scan of each chunk's digit totals, and the
// scatter.  The first pass ranks unstably (its input order is arbitrary); the
// second ranks each key by its combined (second digit, first digit) value,
// which keeps the first pass's order between first digits inside each second
// digit (the order inside one combined value is free: those keys share a
// block group).
// ---------------------------------------------------------------------------
namespace {

constexpr int kSegThreads = 256;
constexpr int kSegIpt = 16;
constexpr uint32_t kSegTile = uint32_t(kSegThreads) * kSegIpt;
constexpr uint32_t kSegMaxBins = 1u << kSegMaxBits;  // combined bins of the second pass
constexpr int kSegBinsPerThread = kSegMaxBins / kSegThreads;
constexpr uint32_t kSegMaxDigits = 1u << kMaxSegDigitBits;
constexpr uint32_t kSegSubCounters = 4;  // LDS words per digit in the count kernel
constexpr int kSegDigitsPerThread = int(kSegMaxDigits / kSegThreads);
static_assert(kSegDigitsPerThread * kSegThreads == int(kSegMaxDigits), "digits spread evenly over the threads");
static_assert(kSegTile == kSegTileItems, "tile size shared with the planner");
static_assert(kSegTile <= 65536 && kSegMaxBins <= 65536, "bin and rank packed in 16 bits each");
static_assert(kSegMaxBins == kSegThreads * kSegBinsPerThread && kSegBinsPerThread * kWave * (kSegThreads / kWave) ==
                  kSegMaxBins, "bins covered by the row-wise scan");

// Counts of tile t, tile-major: counts[t * R + d] (one contiguous run per
// tile, written and read whole).
__device__ __forceinline__ uint64_t seg_count_index(uint64_t t, uint32_t R, uint32_t d)
{
    return t * R + d;
}

// Tile table of the segmented sort, planned on the device from the chunk
// offsets (no host planning, so a captured build replays for any offsets
// within its bounds): chunk c's tiles start at chunk_tiles[c] (exclusive scan
// of ceil(keys / kSegTile)), *ntiles = their total.  *n_ptr == 0 (no keys, or
// offsets a check rejected) plans no tiles.
constexpr int kPlanThreads = 1024;

__global__ __launch_bounds__(kPlanThreads) void k_seg_plan_chunks(const uint64_t* __restrict__ offsets,
                                                                  uint64_t num_chunks,
                                                                  const uint64_t* __restrict__ n_ptr,
                                                                  uint32_t* __restrict__ chunk_tiles,
                                                                  uint64_t* __restrict__ ntiles)
{
    __shared__ uint32_t s_scan[kPlanThreads / kWave];
    const bool any = *n_ptr != 0;
    uint32_t carry = 0;
    for (uint64_t base = 0; base < num_chunks; base += kPlanThreads) {
        const uint64_t c = base + threadIdx.x;
        uint32_t t = 0;
        if (any && c < num_chunks) t = uint32_t((offsets[c + 1] - offsets[c] + kSegTile - 1) / kSegTile);
        uint32_t total;
        const uint32_t ex = block_exclusive_scan<kPlanThreads>(t, s_scan, total);
        if (c < num_chunks) chunk_tiles[c] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) {
        chunk_tiles[num_chunks] = carry;
        *ntiles = carry;
    }
}

// Tile t: the chunk whose tile range holds it (last c with chunk_tiles[c] <= t).
__global__ void k_seg_plan_tiles(const uint64_t* __restrict__ offsets, const uint32_t* __restrict__ chunk_tiles,
                                 uint64_t num_chunks, const uint64_t* __restrict__ ntiles_ptr,
                                 SegTile* __restrict__ tiles)
{
    const uint64_t ntiles = *ntiles_ptr;
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < ntiles; t += stride) {
        uint64_t lo = 0, hi = num_chunks;  // chunk_tiles[lo] <= t < chunk_tiles[hi]
        while (hi - lo > 1) {
            const uint64_t mid = (lo + hi) / 2;
            if (chunk_tiles[mid] <= t) lo = mid;
            else hi = mid;
        }
        const uint64_t start = offsets[lo] + (t - chunk_tiles[lo]) * kSegTile;
        tiles[t] = SegTile{start, uint32_t(min(uint64_t(kSegTile), offsets[lo + 1] - start)), uint32_t(lo)};
    }
}

// Count the pass digit ((key >> shift) & dmask) of every key of each tile.
// The first pass also range-checks the chunk-local keys.  Each digit's counter
// is split over kSegSubCounters LDS words by lane: the second pass's tiles hold
// few distinct digits, and lanes of one wave adding to one word serialise
// (measured 187 -> 140 us for that count on 169M keys).
template <bool FIRST, typename KI>
__global__ __launch_bounds__(kSegThreads) void k_seg_count(const KI* __restrict__ keys,
                                                           const SegTile* __restrict__ tiles,
                                                           const uint64_t* __restrict__ ntiles_ptr, uint32_t shift,
                                                           uint32_t dmask, uint32_t chunk_bits,
                                                           uint32_t* __restrict__ counts,
                                                           unsigned long long* __restrict__ err)
{
    __shared__ uint32_t s_cnt[kSegMaxDigits * kSegSubCounters];
    const uint64_t ntiles = *ntiles_ptr;
    const uint32_t R = dmask + 1;
    bool bad = false;
    // Tiles XCD by XCD: a chunk's neighbouring tiles have adjacent counts in
    // every digit row, which then combine in one L2.
    const TileWalk walk = xcd_tiles(ntiles);
    for (uint64_t t = walk.tile; t < walk.end; t += walk.step) {
        const SegTile tile = tiles[t];
        // The whole tile's keys are in flight before the counters are cleared:
        // positions past the tile's end read its last key (a tile holds at
        // least one), so no load sits behind a branch, and are not counted.
        KI key[kSegIpt];
#pragma unroll
        for (int i = 0; i < kSegIpt; ++i) {
            const uint32_t j = min(uint32_t(i) * kSegThreads + threadIdx.x, tile.count - 1);
            key[i] = __builtin_nontemporal_load(keys + tile.start + j);
        }
        for (uint32_t i = threadIdx.x; i < kSegSubCounters * R; i += kSegThreads) s_cnt[i] = 0;
        __syncthreads();
        const uint32_t sub = threadIdx.x & (kSegSubCounters - 1);
#pragma unroll
        for (int i = 0; i < kSegIpt; ++i) {
            if (uint32_t(i) * kSegThreads + threadIdx.x < tile.count) {
                if constexpr (FIRST) bad |= chunk_bits < 8 * sizeof(KI) && (uint64_t(key[i]) >> chunk_bits) != 0;
                atomicAdd(&s_cnt[(uint32_t(uint64_t(key[i]) >> shift) & dmask) * kSegSubCounters + sub], 1u);
            }
        }
        __syncthreads();
        for (uint32_t d = threadIdx.x; d < R; d += kSegThreads) {
            uint32_t c = 0;
#pragma unroll
            for (uint32_t q = 0; q < kSegSubCounters; ++q) c += s_cnt[d * kSegSubCounters + q];
            counts[seg_count_index(t, R, d)] = c;
        }
        __syncthreads();
    }
    if (FIRST && bad) atomicOr(err, (unsigned long long)kErrKeyRange);
}

// One block per (chunk, 64 digits): lane l of every wave owns digit d = group
// * 64 + l; the chunk's tiles are split into kColWaves consecutive slices, one
// per wave.  Each wave sums its slice (reading 64 adjacent counts of one tile
// per step, kScanAhead tiles in flight), the slice sums are combined in LDS,
// and each wave rewrites its slice with exclusive prefixes; the digit's total
// goes to totals[c * R + d].
constexpr uint32_t kScanAhead = 8;
constexpr int kColWaves = 8;
constexpr int kColThreads = kColWaves * kWave;

__global__ __launch_bounds__(kColThreads) void k_seg_scan_cols(const uint32_t* __restrict__ chunk_tiles,
                                                               uint64_t num_chunks, uint32_t R,
                                                               uint32_t* __restrict__ counts,
                                                               uint32_t* __restrict__ totals)
{
    __shared__ uint32_t s_part[kColWaves][kWave];
    const uint32_t lane = lane_id(), w = threadIdx.x / kWave;
    const uint32_t groups = (R + kWave - 1) / kWave;
    for (uint64_t job = blockIdx.x; job < num_chunks * groups; job += gridDim.x) {
        const uint64_t c = job / groups;
        const uint32_t d = uint32_t(job % groups) * kWave + lane;
        const uint32_t t0 = chunk_tiles[c], nt = chunk_tiles[c + 1] - t0;
        const uint32_t per = (nt + kColWaves - 1) / kColWaves;
        const uint32_t b = min(nt, w * per), e = min(nt, b + per);
        uint32_t* col = counts + uint64_t(t0) * R + d;
        const bool mine = d < R;
        uint32_t sum = 0;
        for (uint32_t i0 = b; i0 < e; i0 += kScanAhead) {
            uint32_t v[kScanAhead];
#pragma unroll
            for (uint32_t k = 0; k < kScanAhead; ++k) v[k] = mine && i0 + k < e ? col[uint64_t(i0 + k) * R] : 0u;
#pragma unroll
            for (uint32_t k = 0; k < kScanAhead; ++k) sum += v[k];
        }
        s_part[w][lane] = sum;
        __syncthreads();
        uint32_t run = 0, total = 0;
#pragma unroll
        for (int i = 0; i < kColWaves; ++i) {
            const uint32_t x = s_part[i][lane];
            run += uint32_t(i) < w ? x : 0u;
            total += x;
        }
        __syncthreads();  // s_part is rewritten by the next job
        if (mine) {
            for (uint32_t i0 = b; i0 < e; i0 += kScanAhead) {
                uint32_t v[kScanAhead];
#pragma unroll
                for (uint32_t k = 0; k < kScanAhead; ++k) v[k] = i0 + k < e ? col[uint64_t(i0 + k) * R] : 0u;
#pragma unroll
                for (uint32_t k = 0; k < kScanAhead; ++k) {
                    if (i0 + k < e) col[uint64_t(i0 + k) * R] = run;
                    run += v[k];
                }
            }
            if (w == 0) totals[c * R + d] = total;
        }
    }
}

// Per chunk: starts[c * R + d] = chunk start + keys of the chunk with a
// smaller digit.
__global__ __launch_bounds__(kSegThreads) void k_seg_scan_digits(const uint64_t* __restrict__ chunk_start,
                                                                 uint64_t num_chunks, uint32_t R,
                                                                 const uint32_t* __restrict__ totals,
                                                                 uint32_t* __restrict__ starts)
{
    __shared__ uint32_t s_scan[kSegThreads / kWave];
    for (uint64_t c = blockIdx.x; c < num_chunks; c += gridDim.x) {
        // kSegDigitsPerThread consecutive digits per thread.
        uint32_t v[kSegDigitsPerThread];
        uint32_t sum = 0;
#pragma unroll
        for (int k = 0; k < kSegDigitsPerThread; ++k) {
            const uint32_t d = threadIdx.x * kSegDigitsPerThread + uint32_t(k);
            v[k] = d < R ? totals[c * R + d] : 0u;
            sum += v[k];
        }
        uint32_t unused;
        uint32_t ex = uint32_t(chunk_start[c]) + block_exclusive_scan<kSegThreads>(sum, s_scan, unused);
#pragma unroll
        for (int k = 0; k < kSegDigitsPerThread; ++k) {
            const uint32_t d = threadIdx.x * kSegDigitsPerThread + uint32_t(k);
            if (d < R) starts[c * R + d] = ex;
            ex += v[k];
        }
    }
}

// The scatter of one pass; keys are written as 32-bit chunk-local keys.
// FIRST: bins = the pass digit.  Else: bins = digit * span + (previous digit -
// the tile's smallest previous digit), span = the tile's range of previous
// digits (the input is sorted on them, so a tile usually holds one or two).
// Output offset of tile t's digit d: starts[c * R + d] plus the row prefix in
// counts.  The first pass needs only R <= kSegMaxDigits bins, which keeps its
// LDS (and so its workgroups per CU) within reach of the later passes' double.
// Tiles are walked XCD by XCD (the grid is a multiple of kXcds), so
// neighbouring tiles, whose runs of one digit are adjacent in the output, are
// written through one L2.
template <bool FIRST, typename KI>
__global__ __launch_bounds__(kSegThreads) void k_seg_scatter(const KI* __restrict__ keys_in,
                                                             uint32_t* __restrict__ keys_out,
                                                             const SegTile* __restrict__ tiles,
                                                             const uint64_t* __restrict__ ntiles_ptr, uint32_t shift,
                                                             uint32_t dmask, uint32_t prev_shift, uint32_t prev_bits,
                                                             const uint32_t* __restrict__ counts,
                                                             const uint32_t* __restrict__ starts)
{
    using KS = KI;  // staged key
    constexpr uint32_t kBins = FIRST ? kSegMaxDigits : kSegMaxBins;
    constexpr int kBinsPerThread = int(kBins / kSegThreads);
    static_assert(kBinsPerThread * kSegThreads == int(kBins), "bins covered by the row-wise scan");
    __shared__ KS s_keys[kSegTile];
    __shared__ uint32_t s_bin[kBins];
    __shared__ uint32_t s_base[kSegMaxDigits];
    __shared__ uint32_t s_scan[kSegThreads / kWave];
    __shared__ uint32_t s_span[2 * (kSegThreads / kWave)];
    const uint64_t ntiles = *ntiles_ptr;
    const uint32_t R = dmask + 1;
    const uint32_t prev_mask = (1u << prev_bits) - 1u;
    const uint32_t w = threadIdx.x / kWave, lane = lane_id();
    const TileWalk walk = xcd_tiles(ntiles);
    for (uint64_t t = walk.tile; t < walk.end; t += walk.step) {
        const SegTile tile = tiles[t];
        KS key[kSegIpt];
        uint32_t br[kSegIpt];  // bin << 16 | rank within the bin
        uint32_t lo = 0xFFFFFFFFu, hi = 0;
        // Every load is issued before any key is used; positions past the
        // tile's end read its last key (no load behind a branch) and are
        // skipped below.
#pragma unroll
        for (int i = 0; i < kSegIpt; ++i) {
            const uint32_t j = min(uint32_t(i) * kSegThreads + threadIdx.x, tile.count - 1);
            key[i] = KS(__builtin_nontemporal_load(keys_in + tile.start + j));
        }
        if constexpr (!FIRST) {
#pragma unroll
            for (int i = 0; i < kSegIpt; ++i) {
                if (uint32_t(i) * kSegThreads + threadIdx.x < tile.count) {
                    const uint32_t p = uint32_t(key[i] >> prev_shift) & prev_mask;
                    lo = min(lo, p);
                    hi = max(hi, p);
                }
            }
        }
        uint32_t pmin = 0, span = 1;
        if constexpr (!FIRST) {
#pragma unroll
            for (int o = kWave / 2; o >= 1; o >>= 1) {
                lo = min(lo, uint32_t(__shfl_xor(int(lo), o, kWave)));
                hi = max(hi, uint32_t(__shfl_xor(int(hi), o, kWave)));
            }
            if (lane == 0) {
                s_span[w] = lo;
                s_span[kSegThreads / kWave + w] = hi;
            }
            __syncthreads();
            lo = s_span[0];
            hi = s_span[kSegThreads / kWave];
#pragma unroll
            for (int i = 1; i < kSegThreads / kWave; ++i) {
                lo = min(lo, s_span[i]);
                hi = max(hi, s_span[kSegThreads / kWave + i]);
            }
            pmin = lo;
            span = hi >= lo ? hi - lo + 1 : 1;
        }
        const uint32_t nbins = R * span;  // FIRST: R; else at most R << prev_bits <= kSegMaxBins
        for (uint32_t i = threadIdx.x; i < nbins; i += kSegThreads) s_bin[i] = 0;
        __syncthreads();
#pragma unroll
        for (int i = 0; i < kSegIpt; ++i) {
            const uint32_t j = uint32_t(i) * kSegThreads + threadIdx.x;
            uint32_t bin = uint32_t(uint64_t(key[i]) >> shift) & dmask;
            if (!FIRST) bin = bin * span + ((uint32_t(key[i] >> prev_shift) & prev_mask) - pmin);
            const uint32_t r = j < tile.count ? atomicAdd(&s_bin[bin], 1u) : 0u;
            br[i] = bin << 16 | r;
        }
        __syncthreads();
        // Exclusive scan of the bins.  Wave w takes bins [w * 64 * kBinsPerThread, ...)
        // as rows of 64 consecutive bins (one per lane, so the LDS accesses
        // are conflict-free), with a carry from row to row and wave to wave;
        // rows past the tile's bins are skipped.
        uint32_t ex[kBinsPerThread];
        uint32_t carry = 0;
#pragma unroll
        for (int q = 0; q < kBinsPerThread; ++q) {
            const uint32_t row0 = (w * kBinsPerThread + uint32_t(q)) * kWave;
            if (row0 >= nbins) break;  // wave-uniform
            const uint32_t b = row0 + lane;
            const uint32_t v = b < nbins ? s_bin[b] : 0u;
            const uint32_t inc = wave_inclusive_scan(v);
            ex[q] = carry + inc - v;
            carry += uint32_t(__builtin_amdgcn_readlane(int(inc), kWave - 1));
        }
        if (lane == 0) s_scan[w] = carry;
        __syncthreads();
        uint32_t base = 0;
#pragma unroll
        for (int i = 0; i < kSegThreads / kWave; ++i) base += uint32_t(i) < w ? s_scan[i] : 0u;
#pragma unroll
        for (int q = 0; q < kBinsPerThread; ++q) {
            const uint32_t row0 = (w * kBinsPerThread + uint32_t(q)) * kWave;
            if (row0 >= nbins) break;
            const uint32_t b = row0 + lane;
            if (b < nbins) s_bin[b] = base + ex[q];
        }
        __syncthreads();
        // Digit d's keys start at its first bin inside the tile.
        for (uint32_t d = threadIdx.x; d < R; d += kSegThreads)
            s_base[d] = starts[uint64_t(tile.chunk) * R + d] + counts[seg_count_index(t, R, d)] -
                        s_bin[d * span];
#pragma unroll
        for (int i = 0; i < kSegIpt; ++i) {
            const uint32_t j = uint32_t(i) * kSegThreads + threadIdx.x;
            if (j < tile.count) s_keys[s_bin[br[i] >> 16] + (br[i] & 0xFFFFu)] = key[i];
        }
        __syncthreads();
        for (uint32_t j = threadIdx.x; j < tile.count; j += kSegThreads) {
            const KS k = s_keys[j];
            keys_out[s_base[uint32_t(uint64_t(k) >> shift) & dmask] + j] = uint32_t(k);
        }
        __syncthreads();
    }
}

}  // namespace

template <typename KI>
int launch_segmented_sort(hipStream_t stream, const KI* keys_in, uint32_t* keys_a, uint32_t* keys_b,
                          const uint64_t* chunk_start, uint64_t num_chunks, const uint64_t* n_ptr, uint64_t n_bound,
                          const SegScratch& sc, uint32_t lo, uint32_t chunk_bits, unsigned long long* err,
                          uint32_t first_bits)
{
    if (chunk_bits <= lo || chunk_bits - lo > kSegMaxBits || chunk_bits > 32 || n_bound == 0) return -1;
    const uint32_t B = chunk_bits - lo;
    // First pass digit width: half the bits unless asked otherwise, and never
    // leaving either pass wider than kMaxSegDigitBits.
    const uint32_t lo_b1 = B > kMaxSegDigitBits ? B - kMaxSegDigitBits : 1u;
    const uint32_t hi_b1 = std::min(B, kMaxSegDigitBits);
    const uint32_t b1 = first_bits ? std::min(hi_b1, std::max(lo_b1, first_bits))
                                   : (B > kMaxSegDigitBits ? (B + 1) / 2 : B);
    const uint32_t b2 = B - b1;                                  // second (0: one pass)
    const uint64_t tile_bound = seg_tile_bound(n_bound, num_chunks);
    const int cus = device_cu_count();
    // A multiple of kXcds, as the scatter's tile walk assumes.
    const int grid = int(std::max<uint64_t>(
        kXcds, std::min<uint64_t>((tile_bound + kXcds - 1) / kXcds * kXcds, uint64_t(cus) * 64) / kXcds * kXcds));
    const int chunk_grid = int(std::max<uint64_t>(1, std::min<uint64_t>(num_chunks, uint64_t(cus) * 8)));
    {
        TimedLaunch t(kKSortScan, stream);
        hipLaunchKernelGGL(k_seg_plan_chunks, dim3(1), dim3(kPlanThreads), 0, stream, chunk_start, num_chunks, n_ptr,
                           sc.chunk_tiles, sc.ntiles);
        hipLaunchKernelGGL(k_seg_plan_tiles, dim3(int(std::max<uint64_t>(1, std::min<uint64_t>((tile_bound + 255) / 256,
                                                                                           uint64_t(cus) * 8)))),
                           dim3(256), 0, stream, chunk_start, sc.chunk_tiles, num_chunks, sc.ntiles, sc.tiles);
    }
    // The first pass reads the caller's keys (KI), the second the first's output.
    auto pass = [&](bool first, const void* src, uint32_t* dst, uint32_t shift, uint32_t bits) {
        const uint32_t mask = (1u << bits) - 1u, R = mask + 1;
        const uint64_t col_jobs = num_chunks * ((R + kWave - 1) / kWave);
        const int col_grid = int(std::max<uint64_t>(1, std::min<uint64_t>(col_jobs, uint64_t(cus) * 16)));
        {
            TimedLaunch t(kKSegCount, stream);
            if (first)
                hipLaunchKernelGGL((k_seg_count<true, KI>), dim3(grid), dim3(kSegThreads), 0, stream,
                                   static_cast<const KI*>(src), sc.tiles, sc.ntiles, shift, mask,
                                   chunk_bits, sc.counts, err);
            else
                hipLaunchKernelGGL((k_seg_count<false, uint32_t>), dim3(grid), dim3(kSegThreads), 0, stream,
                                   static_cast<const uint32_t*>(src), sc.tiles, sc.ntiles, shift,
                                   mask, chunk_bits, sc.counts, err);
        }
        {
            TimedLaunch t(kKSortScan, stream);
            hipLaunchKernelGGL(k_seg_scan_cols, dim3(col_grid), dim3(kColThreads), 0, stream, sc.chunk_tiles,
                               num_chunks, R,
                               sc.counts, sc.totals);
            hipLaunchKernelGGL(k_seg_scan_digits, dim3(chunk_grid), dim3(kSegThreads), 0, stream, chunk_start,
                               num_chunks, R, sc.totals, sc.starts);
        }
        TimedLaunch t(kKSegScatter, stream);
        if (first)
            hipLaunchKernelGGL((k_seg_scatter<true, KI>), dim3(grid), dim3(kSegThreads), 0, stream,
                               static_cast<const KI*>(src), dst, sc.tiles, sc.ntiles, shift, mask, 0u, 0u, sc.counts,
                               sc.starts);
        else
            hipLaunchKernelGGL((k_seg_scatter<false, uint32_t>), dim3(grid), dim3(kSegThreads), 0, stream,
                               static_cast<const uint32_t*>(src), dst, sc.tiles, sc.ntiles, shift, mask, lo, b1,
                               sc.counts, sc.starts);
    };
    pass(true, keys_in, keys_a, lo, b1);
    if (b2 == 0) return 0;
    pass(false, keys_a, keys_b, lo + b1, b2);
    return 1;
}

template int launch_segmented_sort<uint64_t>(hipStream_t, const uint64_t*, uint32_t*, uint32_t*, const uint64_t*,
                                             uint64_t, const uint64_t*, uint64_t, const SegScratch&, uint32_t,
                                             uint32_t, unsigned long long*, uint32_t);
template int launch_segmented_sort<uint32_t>(hipStream_t, const uint32_t*, uint32_t*, uint32_t*, const uint64_t*,
                                             uint64_t, const uint64_t*, uint64_t, const SegScratch&, uint32_t,
                                             uint32_t, unsigned long long*, uint32_t);

// ---------------------------------------------------------------------------
// Top digit first (plain builds of 64-bit keys).  One pass on the top digit
// [top_shift, top_shift + top_bits) partitions the keys into 2^top_bits
// Morton-ordered buckets and writes each key as a 32-bit key local to its
// bucket (its bits below top_shift); the segmented sort then sorts the bits
// [lo, top_shift) inside each bucket, reading and writing those 32-bit keys
// (the leaf reduce restores the bucket index from the bucket starts).  Against
// an LSD sort of the same bits it moves half the bytes in the top pass's writes
// and in every read and write after it.
// ---------------------------------------------------------------------------
namespace {

// starts[d] = keys with a smaller top digit (d < R); starts[R] = n.
__global__ __launch_bounds__(kRadix) void k_digit_starts(const uint32_t* __restrict__ hist, uint32_t R,
                                                         const uint64_t* __restrict__ n_ptr,
                                                         uint64_t* __restrict__ starts)
{
    __shared__ uint32_t s_scan[kRadix / kWave];
    const uint32_t d = threadIdx.x;
    uint32_t total;
    const uint32_t ex = block_exclusive_scan<kRadix>(d < R ? hist[d] : 0u, s_scan, total);
    if (d < R) starts[d] = ex;
    if (d == 0) starts[R] = *n_ptr;
}

}  // namespace

int launch_top_digit_sort(hipStream_t stream, const uint64_t* keys_in, uint32_t* local, uint32_t* keys_a,
                          uint32_t* keys_b, int ipt, uint32_t top_shift, uint32_t top_bits, uint32_t lo,
                          const uint64_t* n_ptr, uint64_t n_bound, uint32_t* zero, uint32_t* counts,
                          unsigned long long* key_or, uint64_t* bucket_start, const SegScratch& sc,
                          unsigned long long* err)
{
    if (top_bits < 1 || top_bits > uint32_t(kMaxRadixBits) || top_shift > 32 || top_shift <= lo ||
        top_shift - lo > kSegMaxBits || n_bound == 0)
        return -1;
    const uint32_t dmask = (1u << top_bits) - 1u, R = dmask + 1;
    HistPlan hp{};
    hp.passes = 1;
    hp.shift[0] = top_shift;
    hp.mask[0] = dmask;
    auto run = [&](auto ipt_tag) {
        constexpr int I = decltype(ipt_tag)::value;
        const uint64_t stride = sort_tiles(n_bound, I);
        const int cus = device_cu_count();
        const uint32_t scatter_grid = uint32_t(std::max<uint64_t>(
            kXcds, std::min<uint64_t>((stride + kXcds - 1) / kXcds * kXcds, uint64_t(cus) * 64) / kXcds * kXcds));
        {
            TimedLaunch t(kKHistogram, stream);
            hipLaunchKernelGGL((k_histogram<uint64_t, I>), dim3(histogram_grid<uint64_t, I>(stride)),
                               dim3(kSortThreads), 0, stream, keys_in, n_ptr, hp, zero, key_or, counts, stride,
                               ChunkPrefix{});
        }
        {
            TimedLaunch t(kKSortScan, stream);
            if (stride <= kSmallScanTiles)
                hipLaunchKernelGGL(k_sort_scan<kWave>, dim3(R), dim3(kWave), 0, stream, counts, zero, n_ptr, stride,
                                   uint64_t(kSortThreads) * I);
            else
                hipLaunchKernelGGL(k_sort_scan<kScanThreads>, dim3(R), dim3(kScanThreads), 0, stream, counts, zero,
                                   n_ptr, stride, uint64_t(kSortThreads) * I);
            hipLaunchKernelGGL(k_digit_starts, dim3(1), dim3(kRadix), 0, stream, zero, R, n_ptr, bucket_start);
        }
        // Without the LDS that holds the 64-bit LSD pass to two workgroups per
        // CU: this pass writes half the bytes and measured faster at four
        // (207 against 243 us on 70M keys).
        TimedLaunch t(kKSortScatterTop, stream);
        hipLaunchKernelGGL((k_sort_scatter<uint64_t, false, kRankUnstable, I, kSortThreads, uint32_t>),
                           dim3(scatter_grid), dim3(kSortThreads), 0, stream, keys_in, local, nullptr,
                           nullptr, n_ptr, top_shift, dmask, counts, stride, ChunkPrefix{}, 0u, 0u, nullptr, 64u);
    };
    if (ipt == kSortIptLarge) run(std::integral_constant<int, kSortIptLarge>{});
    else run(std::integral_constant<int, kSortIptSmall>{});
    return launch_segmented_sort<uint32_t>(stream, local, keys_a, keys_b, bucket_start, R, n_ptr, n_bound, sc, lo,
                                           top_shift, err, 0);
}

}  // namespace svdag
