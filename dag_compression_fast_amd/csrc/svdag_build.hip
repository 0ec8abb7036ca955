// svdag_build.hip -- gfx950 kernels of the bottom-up DAG build.
//
// Reference path (src/DAGCompression.cu):
//   SelectUnique + leaf ReduceByKey  :1035-1068  -> k_leaf_reduce / k_leaf_compact
//       (one pass over the keys grouped by code >> 15: duplicates OR into the
//       same 4^3 mask, so no separate unique step is needed)
//   SortLevel(Leaves)               :1070-1071  -> k_table_clear / k_leaf_insert
//       (hash table keyed by the mask; the first inserter of a mask appends it
//       to the unique list) + a key/value radix sort of that list +
//       k_rank_of_slot
//   ExtractLevel + ComputeHashes    :916-974, :827-833 -> k_level_count /
//       k_level_build (parent masks, child ranks, Murmur hash, insert into the
//       level's dedup table)
//   SortLevel(level)                :864-914    -> k_extract_count /
//       k_level_extract (ordered linear-probing table: home slot = top bits of
//       the hash, so sorting each probe cluster orders the whole table; rank =
//       occupied slots before it)
//   CreateFinalDAG / WriteTo        :834-861, :1550-1610 -> k_final_count /
//       k_final_positions / k_final_write
// Every compaction is reduce-then-scan: a count kernel writes one count per
// tile, k_scan_counts (one workgroup) turns them into offsets, and the main
// kernel handles tile blockIdx.x + k * gridDim.x.  No workgroup waits on
// another; element counts come from device memory, so the build runs without
// host round trips.
#include "svdag_device.hpp"
#include "svdag_internal.hpp"

#include <algorithm>
#include <type_traits>

namespace svdag {

namespace {

constexpr int kThreads = 256;
constexpr int kScanThreads = 1024;
constexpr int kScanItems = 4;

__host__ int grid_for(uint64_t tiles, int per_cu)
{
    const uint64_t cap = uint64_t(device_cu_count()) * per_cu;
    return int(std::max<uint64_t>(1, std::min<uint64_t>(tiles, cap)));
}

__global__ void k_fill_u64(uint64_t* __restrict__ p, uint64_t count, uint64_t value, uint64_t set_at,
                           uint64_t set_value)
{
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride)
        p[i] = i == set_at ? set_value : value;
}

// Exclusive scan of counts[0, ntiles) into offsets, ntiles = ceil(*items / per_tile);
// the grand total goes to *total_out when given.  One workgroup.
__global__ __launch_bounds__(kScanThreads) void k_scan_counts(const uint32_t* __restrict__ counts,
                                                              uint64_t* __restrict__ offsets,
                                                              const uint64_t* __restrict__ items, uint64_t per_tile,
                                                              uint64_t* __restrict__ total_out, uint64_t items_host)
{
    __shared__ uint64_t s_scan[kScanThreads / kWave];
    const uint64_t ntiles = ((items ? *items : items_host) + per_tile - 1) / per_tile;
    uint64_t running = 0;
    // kScanItems consecutive counts per thread per block scan.
    for (uint64_t chunk = 0; chunk < ntiles; chunk += uint64_t(kScanThreads) * kScanItems) {
        const uint64_t i0 = chunk + uint64_t(threadIdx.x) * kScanItems;
        uint32_t v[kScanItems];
        uint64_t sum = 0;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            v[k] = i0 + k < ntiles ? counts[i0 + k] : 0u;
            sum += v[k];
        }
        uint64_t total;
        uint64_t run = running + block_exclusive_scan64<kScanThreads>(sum, s_scan, total);
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            if (i0 + k < ntiles) offsets[i0 + k] = run;
            run += v[k];
        }
        running += total;
    }
    if (total_out && threadIdx.x == 0) *total_out = running;
}

// ---------------------------------------------------------------------------
// Leaf reduce: keys sorted by code >> (6 + r), with r = min(9, 3 * depth - 6)
// -> (code >> 6, OR of 1 << (code & 63)) per 4^3 block, in code order.
// A run of keys with equal code >> (6 + r) (a "block group", at most 2^r
// leaves) is reduced in a wave-private LDS table of 2^r words indexed by the
// leaf's low r bits; the keys of a group may come in any order, so the sort
// skips the low r bits of the leaf code as well as the 6 voxel bits.
// Wave work unit: a chunk of kPackTileItems consecutive keys.  A wave owns the
// groups whose first key lies in its chunk (reading on past the chunk to
// finish its last one) and writes their leaves into its own region of
// kPackTileItems + 512 entries: groups inside the chunk have at most one leaf
// per key, the last group at most 2^r leaves.  k_leaf_compact then moves the
// regions together.
// ---------------------------------------------------------------------------
constexpr uint32_t kReduceChunk = uint32_t(kPackTileItems);
constexpr uint32_t kReduceTable = 1u << kLeafGroupBits;
constexpr uint32_t kReduceRegion = uint32_t(kLeafRegionItems);
static_assert(kReduceRegion == kReduceChunk + kReduceTable, "region bound");
// Groups of 64 keys loaded together: 2 KiB of keys per wave in flight.
template <typename K>
constexpr int kReduceGroups = int(2048 / (kWave * sizeof(K)));
constexpr int kReduceWaves = kThreads / kWave;
static_assert(kReduceTable == 8 * kWave, "one lane holds the 8 leaves of one parent");

// Value of lane l (a wave-uniform l), as a scalar.
__device__ __forceinline__ uint64_t readlane64(uint64_t v, uint32_t l)
{
    const uint32_t lo = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(v)), int(l)));
    const uint32_t hi = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(v >> 32)), int(l)));
    return uint64_t(hi) << 32 | lo;
}

// G groups of 64 raw keys from `base`; indices past the end read
// the last key (n >= 1), and the caller ignores those lanes.  No use of the
// values here, so the loads stay in flight until the keys are needed.
template <typename K, int G, typename KC>
__device__ __forceinline__ void load_groups(const K* __restrict__ sorted, uint64_t base, uint64_t n,
                                            KC (&k)[G])
{
    const uint32_t lane = lane_id();
#pragma unroll
    for (int g = 0; g < G; ++g)  // read once: non-temporal
        k[g] = KC(__builtin_nontemporal_load(sorted + min(base + g * kWave + lane, n - 1)));
}

// Lane l's value of v (wave-uniform result), for 32- and 64-bit values.
__device__ __forceinline__ uint32_t readlane_of(uint32_t v, uint32_t l)
{
    return uint32_t(__builtin_amdgcn_readlane(int(v), int(l)));
}
__device__ __forceinline__ uint64_t readlane_of(uint64_t v, uint32_t l) { return readlane64(v, l); }
__device__ __forceinline__ uint32_t shfl_up_of(uint32_t v) { return uint32_t(__shfl_up(int(v), 1, kWave)); }
__device__ __forceinline__ uint64_t shfl_up_of(uint64_t v) { return __shfl_up(v, 1, kWave); }

// Bucket of position p among Morton-ordered buckets with starts start(0..num)
// (start(num) > p): the last b >= lo with start(b) <= p (empty buckets share
// their start with the next one).
template <typename Start>
__device__ __forceinline__ uint64_t bucket_of(const Start& start, uint64_t lo, uint64_t num, uint64_t p)
{
    uint64_t hi = num;
    while (hi - lo > 1) {
        const uint64_t mid = (lo + hi) / 2;
        if (start(mid) <= p) lo = mid;
        else hi = mid;
    }
    return lo;
}

constexpr uint32_t kLdsBuckets = 512;  // bucket tables up to this size are staged in LDS

// Where the leaf reduce finds bucket starts: no buckets (keys are codes), a
// copy in LDS (small tables), or the table in global memory.
enum BucketMode { kBucketsNone = 0, kBucketsLds = 1, kBucketsGlobal = 2 };

// BM != kBucketsNone: 32-bit bucket-local keys (see launch_leaf_reduce); each
// wave keeps the bucket of its current position and its end, so a batch inside
// one bucket takes the prefix from a register; batches that cross a bucket end
// (a few per bucket) search per lane.
template <typename K, int BM>
__global__ __launch_bounds__(kThreads) void k_leaf_reduce(const K* __restrict__ sorted,
                                                          const uint64_t* __restrict__ n_ptr, uint64_t code_mask,
                                                          uint32_t r, uint64_t* __restrict__ reg_code,
                                                          uint64_t* __restrict__ reg_mask,
                                                          uint32_t* __restrict__ counts,
                                                          uint64_t* __restrict__ next_count,
                                                          unsigned long long* __restrict__ err, ChunkPrefix bk)
{
    __shared__ uint64_t s_tab[kReduceWaves][kReduceTable];
    __shared__ uint64_t s_parents[kReduceWaves];
    constexpr int G = kReduceGroups<K>;
    constexpr bool BUCKETED = BM != kBucketsNone;
    // Keys and group indices in registers: 32-bit for 32-bit keys (a group
    // index then has at most 32 - 15 local bits and the bucket's 12 above
    // them), 64-bit otherwise.
    using KC = std::conditional_t<sizeof(K) == 4, uint32_t, uint64_t>;
    // Bucket starts (positions < 2^30) staged in LDS when the table is small.
    __shared__ uint32_t s_bstart[BM == kBucketsLds ? kLdsBuckets + 1 : 1];
    if constexpr (BM == kBucketsLds) {
        for (uint32_t i = threadIdx.x; i <= bk.num_chunks; i += kThreads) s_bstart[i] = uint32_t(bk.offsets[i]);
        __syncthreads();
    }
    auto start_of = [&](uint64_t b) -> uint64_t {
        if constexpr (BM == kBucketsLds) return s_bstart[b];
        else return bk.offsets[b];
    };
    const uint32_t lane = lane_id();
    // The wave index as a scalar, so that the chunk loop and the group
    // bookkeeping below compile to scalar control flow.
    const uint32_t w = uint32_t(__builtin_amdgcn_readfirstlane(int(threadIdx.x / kWave)));
    uint64_t* const tab = s_tab[w];
#pragma unroll
    for (int e = 0; e < 8; ++e) tab[e * kWave + lane] = 0ull;
    __builtin_amdgcn_wave_barrier();
    const uint64_t n = *n_ptr;
    const uint64_t nchunks = (n + kReduceChunk - 1) / kReduceChunk;
    const uint32_t sh = 6 + r;
    const uint32_t low_mask = (1u << r) - 1u;
    uint32_t parent_lanes = 0;  // per lane: parents counted by the flushes
    bool overflow = false;

    for (uint64_t chunk = uint64_t(blockIdx.x) * kReduceWaves + w; chunk < nchunks;
         chunk += uint64_t(gridDim.x) * kReduceWaves) {
        const uint64_t c0 = chunk * kReduceChunk;
        const uint64_t c1 = min(c0 + kReduceChunk, n);
        uint64_t* const oc = reg_code + chunk * kReduceRegion;
        uint64_t* const om = reg_mask + chunk * kReduceRegion;
        uint32_t written = 0;
        // Bucket of the chunk's current batch and that bucket's end.
        uint64_t cur = 0, cur_end = ~0ull;
        if constexpr (BUCKETED) {
            cur = bucket_of(start_of, 0, bk.num_chunks, c0 > 0 ? c0 - 1 : 0);
            cur_end = start_of(cur + 1);
        }
        // Group indices (code >> sh) of the batch of keys v at bb: the key's own
        // bits above sh, with its bucket above those (bucketed keys).
        const uint32_t gsh = BUCKETED ? bk.chunk_bits - sh : 0u;  // chunk_bits > sh
        auto group_of = [&](uint64_t bb, const KC (&v)[G], KC (&out)[G]) {
            if constexpr (BUCKETED) {
                const uint64_t last = min(bb + uint64_t(G) * kWave - 1, n - 1);
                if (last < cur_end) {
                    const KC pre = KC(cur) << gsh;
#pragma unroll
                    for (int g = 0; g < G; ++g) out[g] = pre | (v[g] >> sh);
                } else {
#pragma unroll
                    for (int g = 0; g < G; ++g) {
                        const uint64_t p = min(bb + uint64_t(g) * kWave + lane, n - 1);
                        out[g] = (KC(bucket_of(start_of, cur, bk.num_chunks, p)) << gsh) | (v[g] >> sh);
                    }
                    cur = bucket_of(start_of, cur, bk.num_chunks, last);
                    cur_end = start_of(cur + 1);
                }
            } else {
#pragma unroll
                for (int g = 0; g < G; ++g) out[g] = v[g] >> sh;
            }
        };
        // Group of the key before the chunk (~0 never equals a group index).
        KC prev_b = KC(~0ull);
        if (c0 > 0) {
            const KC pk = KC(sorted[c0 - 1]) & KC(code_mask);
            if constexpr (BUCKETED) prev_b = (KC(cur) << gsh) | (pk >> sh);  // cur: the bucket of c0 - 1
            else prev_b = pk >> sh;
        }
        bool open = false;  // a group owned by this wave is being accumulated
        KC cur_b = 0;
        bool done = false;

        // Writes the open group's leaves in order and clears the table.  Row e
        // of the table is read by lanes 0..63 (leaves 64e + lane), so a leaf's
        // output position is the count of nonzero rows' ballots before it;
        // the 8 leaves of one parent are 8 consecutive lanes of one row.
        auto flush = [&]() {
            __builtin_amdgcn_wave_barrier();
            // A full table fits the region for valid bounds: the groups before
            // the last hold at most one leaf per chunk key, so written <= 4095.
            // Past that, an internal bound failed: reported, nothing written.
            const bool room = written + kReduceTable <= kReduceRegion;
            if (!room) overflow = true;
            // Row by row, with the per-row bookkeeping in vector registers:
            // this runs once per group and its scalar work had been about half
            // of the kernel's scalar instructions (the scalar unit is its
            // limit).  Lane l of row e counts parent 8e + l/8 when l % 8 == 0
            // and one of its 8 leaves is present.
            uint32_t pos = written;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const uint64_t m = tab[e * kWave + lane];
                tab[e * kWave + lane] = 0ull;
                const uint64_t B = __ballot(m != 0ull);
                parent_lanes += ((lane & 7u) == 0u && ((B >> (lane & 56u)) & 0xFFull) != 0ull) ? 1u : 0u;
                if (m && room) {
                    const uint32_t p = pos + uint32_t(__popcll(B & lanemask_lt()));
                    oc[p] = (uint64_t(cur_b) << r) | uint64_t(e * kWave + lane);
                    om[p] = m;
                }
                pos += uint32_t(__popcll(B));
            }
            if (room) written = pos;
            __builtin_amdgcn_wave_barrier();
            open = false;
        };

        KC k[G], nxt[G], gb[G];
        load_groups(sorted, c0, n, k);
        // No load may be pending at the loop head (see the wait at its end).
        __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
#pragma unroll
        for (int g = 0; g < G; ++g) k[g] &= KC(code_mask);
        group_of(c0, k, gb);
        for (uint64_t base = c0; !done; base += uint64_t(G) * kWave) {
            // The next batch of the chunk is in flight while this one is reduced.
            const uint64_t next = base + uint64_t(G) * kWave;
            const bool prefetch = next < c1;
            if (prefetch) load_groups(sorted, next, n, nxt);
            // Group starts of the whole batch first: a lane starts a group when
            // its group index differs from the previous lane's (lane 0: from
            // the previous batch group's last lane).  A shuffle and a ballot
            // for every batch group: the shortcut of comparing the first and
            // last lanes (no start inside when they agree) cost more scalar
            // work and a branch than it saved.  Lanes past the end repeat the
            // last key: they never start a group, and OR-ing their bit again is
            // harmless.
            uint64_t H[G];
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const KC b = gb[g];
                KC pb = shfl_up_of(b);
                if (lane == 0) pb = prev_b;
                H[g] = __ballot(b != pb);
                prev_b = readlane_of(b, kWave - 1);
            }
            // A batch that ends before both the chunk's and the keys' end (all
            // but the last of a chunk) cannot end the wave's work: FAST drops
            // the per-group end checks (scalar work; the scalar unit is this
            // kernel's limit).
            auto reduce_batch = [&](auto fast_tag) {
                constexpr bool FAST = decltype(fast_tag)::value;
#pragma unroll
                for (int g = 0; g < G; ++g) {
                    if (!FAST && done) break;
                    const uint64_t gi = base + uint64_t(g) * kWave;  // index of lane 0's key
                    const KC key = k[g];
                    uint64_t* const word = &tab[uint32_t(key >> 6) & low_mask];
                    const uint64_t bit = 1ull << (uint32_t(key) & 63u);
                    uint64_t heads = H[g];
                    const bool keys_end = !FAST && gi + kWave >= n;  // the last key is in this batch group
                    // Segments [lo, next start), then [lo, 64): lanes before the first
                    // start continue the open group, or one owned by an earlier wave
                    // when none is open.  One flush site (the open group ends at a
                    // start, or at the keys' end): every call site is a copy of it.
                    uint32_t lo = 0;
                    while (true) {
                        const uint32_t hi = heads ? uint32_t(__ffsll((long long)heads) - 1) : uint32_t(kWave);
                        if (open && lane >= lo && lane < hi) atomicOr((unsigned long long*)word, bit);
                        const bool last = hi == uint32_t(kWave);
                        if (open && (!last || keys_end)) flush();
                        if (last) {
                            // The keys end here, or no group starts in the chunk.
                            if (!FAST && (keys_end || (!open && gi + kWave >= c1))) done = true;
                            break;
                        }
                        if (!FAST && gi + hi >= c1) {
                            done = true;
                            break;
                        }
                        open = true;
                        cur_b = readlane_of(gb[g], hi);
                        heads &= heads - 1ull;
                        lo = hi;
                    }
                }
            };
            if (next < c1 && next < n) reduce_batch(std::true_type{});
            else reduce_batch(std::false_type{});
            if (done) break;
            if (!prefetch) load_groups(sorted, next, n, nxt);
            // Settle every load here (vmcnt 0, other counters left alone): the
            // next batch's keys are then complete at the loop head, so the
            // batch's first uses do not wait behind the prefetch issued there.
            __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
#pragma unroll
            for (int g = 0; g < G; ++g) k[g] = nxt[g] & KC(code_mask);
            group_of(next, k, gb);
        }
        if (lane == 0) counts[chunk] = written;
    }
    uint64_t parents = parent_lanes;
#pragma unroll
    for (int d = kWave / 2; d >= 1; d >>= 1) parents += uint64_t(uint32_t(__shfl_xor(int(uint32_t(parents)), d, kWave)));
    if (lane == 0) s_parents[w] = parents;
    if (overflow && lane == 0) atomicOr(err, (unsigned long long)kErrInternal);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t t = 0;
        for (int i = 0; i < kReduceWaves; ++i) t += s_parents[i];
        if (t) atomicAdd((unsigned long long*)next_count, (unsigned long long)t);
    }
}

// Region of chunk c (counts[c] leaves) -> lp_code / lp_mask at offsets[c].
// One wave per chunk.  A total beyond the capacity (distinct leaves never
// exceed it) is reported, not written.
__global__ __launch_bounds__(kThreads) void k_leaf_compact(const uint64_t* __restrict__ reg_code,
                                                           const uint64_t* __restrict__ reg_mask,
                                                           const uint32_t* __restrict__ counts,
                                                           const uint64_t* __restrict__ offsets,
                                                           const uint64_t* __restrict__ n_ptr,
                                                           uint64_t* __restrict__ lp_code,
                                                           uint64_t* __restrict__ lp_mask, uint64_t capacity,
                                                           unsigned long long* __restrict__ err)
{
    const uint32_t lane = lane_id();
    const uint64_t nchunks = (*n_ptr + kReduceChunk - 1) / kReduceChunk;
    for (uint64_t chunk = uint64_t(blockIdx.x) * kReduceWaves + threadIdx.x / kWave; chunk < nchunks;
         chunk += uint64_t(gridDim.x) * kReduceWaves) {
        const uint32_t cnt = counts[chunk];
        const uint64_t dst = offsets[chunk];
        const uint64_t src = chunk * kReduceRegion;
        if (dst + cnt > capacity) {
            if (lane == 0) atomicOr(err, (unsigned long long)kErrInternal);
            continue;
        }
        for (uint32_t j = lane; j < cnt; j += kWave) {
            lp_code[dst + j] = reg_code[src + j];
            lp_mask[dst + j] = reg_mask[src + j];
        }
    }
}

// Clears slots [0, 2^lg + overflow) with lg = max(8, ceil(log2(count << load_shift))).
__global__ void k_table_clear(uint64_t* __restrict__ keys, const uint64_t* __restrict__ count_ptr,
                              uint32_t* __restrict__ log2_out, uint32_t load_shift, uint32_t overflow)
{
    const uint64_t cnt = *count_ptr;
    const uint32_t need = ceil_log2_u64((cnt ? cnt : 1ull) << load_shift);
    const uint32_t lg = need > 8u ? need : 8u;
    const uint64_t total = (1ull << lg) + overflow;
    if (blockIdx.x == 0 && threadIdx.x == 0) *log2_out = lg;
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += stride) keys[i] = 0;
}

// Inserts key h (never 0) into a wrap-around table of 2^lg slots; `won` tells
// whether this call placed the key.  A probe never exceeds the table size; a
// full table reports kErrTableOverflow.
__device__ __forceinline__ uint32_t hashed_insert(uint64_t* __restrict__ tk, uint32_t lg, uint64_t h, bool& won,
                                                  unsigned long long* err)
{
    won = false;
    const uint64_t wrap = (1ull << lg) - 1;
    uint64_t slot = fmix64(h) >> (64 - lg);
    for (uint64_t probe = 0; probe <= wrap; ++probe, slot = (slot + 1) & wrap) {
        const uint64_t cur = probe_load(&tk[slot]);
        if (cur == h) return uint32_t(slot);
        if (cur == 0) {
            const uint64_t prev = atomicCAS((unsigned long long*)&tk[slot], 0ull, (unsigned long long)h);
            if (prev == 0) {
                won = true;
                return uint32_t(slot);
            }
            if (prev == h) return uint32_t(slot);
        }
    }
    atomicOr(err, (unsigned long long)kErrTableOverflow);
    return 0;
}

// Per-wave election among lanes holding equal 64-bit keys: every valid lane
// writes its lane index into the wave's LDS bucket (kElectBuckets words) for
// its key, one write wins; a lane whose bucket winner holds the same key is a
// follower of that winner and returns it, other lanes return themselves.
// Every lane of the wave must call it.
constexpr uint32_t kElectBuckets = 128;

__device__ __forceinline__ uint32_t elect_leader(uint64_t key, bool valid, uint32_t* buckets, bool& follower)
{
    const uint32_t lane = lane_id();
    const uint32_t bucket = uint32_t(fmix64(key)) & (kElectBuckets - 1);
    if (valid) buckets[bucket] = lane;
    __builtin_amdgcn_wave_barrier();
    const uint32_t winner = valid ? buckets[bucket] : lane;
    __builtin_amdgcn_wave_barrier();
    const uint64_t winner_key = __shfl(key, int(winner), kWave);
    follower = valid && winner != lane && winner_key == key;
    return follower ? winner : lane;
}

// Every leaf position's mask (never 0) enters the table and records its slot.
// A mask shared by many leaves (all-ones inside a solid, the same few patterns
// along a flat surface) would otherwise turn into a stream of probes and
// atomics on one table slot: a CU whose L1 still holds the slot's old 0
// settles each such probe with a CAS on the same L2 line.  So each wave first
// elects one lane per distinct mask: every lane writes its lane index into a
// per-wave LDS bucket chosen by its mask (one write wins), and a lane whose
// bucket winner holds the same mask takes the winner's slot instead of probing.
// Lanes that lose their bucket to another mask probe themselves.
constexpr int kInsertIpt = 8;

__global__ __launch_bounds__(kThreads) void k_leaf_insert(const uint64_t* __restrict__ lp_mask,
                                                          const uint64_t* __restrict__ count_ptr,
                                                          uint64_t* __restrict__ tk,
                                                          const uint32_t* __restrict__ log2_ptr,
                                                          uint32_t* __restrict__ lp_slot,
                                                          unsigned long long* __restrict__ err)
{
    __shared__ uint32_t s_elect[kThreads / kWave][kElectBuckets];
    const uint64_t n = *count_ptr;
    const uint32_t lg = *log2_ptr;
    const uint32_t lane = lane_id();
    uint32_t* const elect = s_elect[threadIdx.x / kWave];
    constexpr uint64_t per_round = uint64_t(kThreads) * kInsertIpt;
    // The loop bound is wave-uniform, so every lane of a wave takes part in the
    // shuffles below.
    for (uint64_t r0 = uint64_t(blockIdx.x) * per_round + (threadIdx.x & ~uint32_t(kWave - 1)); r0 < n;
         r0 += uint64_t(gridDim.x) * per_round) {
        // The round's masks are all in flight before the first insert (indices
        // past the end read the last mask and are not inserted).
        uint64_t masks[kInsertIpt];
#pragma unroll
        for (int j = 0; j < kInsertIpt; ++j) masks[j] = lp_mask[min(r0 + uint64_t(j) * kThreads + lane, n - 1)];
#pragma unroll
        for (int j = 0; j < kInsertIpt; ++j) {
            const uint64_t q = r0 + uint64_t(j) * kThreads + lane;
            const bool valid = q < n;
            const uint64_t mask = valid ? masks[j] : 0ull;
            bool follower;
            const uint32_t leader = elect_leader(mask, valid, elect, follower);
            uint32_t slot = 0;
            if (valid && !follower) {
                bool won;
                slot = hashed_insert(tk, lg, mask, won, err);
            }
            slot = uint32_t(__shfl(int(slot), int(leader), kWave));
            if (valid) lp_slot[q] = slot;
        }
    }
}

// Unique list of a hashed table: its occupied slots (key, slot index), by
// count / scan / write over tiles of kListTile slots.
constexpr int kListIpt = 16;
constexpr uint32_t kListTile = kThreads * kListIpt;

__global__ __launch_bounds__(kThreads) void k_slot_count(const uint64_t* __restrict__ tk,
                                                         const uint32_t* __restrict__ log2_ptr,
                                                         uint32_t* __restrict__ counts, uint64_t* __restrict__ slots_out)
{
    __shared__ uint64_t s_red[kThreads / kWave];
    const uint64_t total = 1ull << *log2_ptr;
    const uint64_t ntiles = (total + kListTile - 1) / kListTile;
    if (blockIdx.x == 0 && threadIdx.x == 0) *slots_out = total;
    for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        uint32_t occ = 0;
        // Unconditional loads (indices past the end read the last slot and are
        // not counted), so that all are in flight at once.
        uint64_t v[kListIpt];
#pragma unroll
        for (int j = 0; j < kListIpt; ++j) v[j] = tk[min(tile * kListTile + uint64_t(j) * kThreads + threadIdx.x, total - 1)];
#pragma unroll
        for (int j = 0; j < kListIpt; ++j) {
            const uint64_t i = tile * kListTile + uint64_t(j) * kThreads + threadIdx.x;
            occ += (i < total && v[j] != 0) ? 1u : 0u;
        }
        const uint64_t cnt = block_reduce_sum<kThreads>(occ, s_red);
        if (threadIdx.x == 0) counts[tile] = uint32_t(cnt);
    }
}

__global__ __launch_bounds__(kThreads) void k_slot_list(const uint64_t* __restrict__ tk,
                                                        const uint32_t* __restrict__ log2_ptr,
                                                        const uint64_t* __restrict__ offsets,
                                                        uint64_t* __restrict__ list_keys,
                                                        uint32_t* __restrict__ list_slots)
{
    __shared__ uint32_t s_scan[kThreads / kWave];
    const uint64_t total = 1ull << *log2_ptr;
    const uint64_t ntiles = (total + kListTile - 1) / kListTile;
    for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        uint64_t key[kListIpt];
        uint32_t occ = 0;
#pragma unroll
        for (int j = 0; j < kListIpt; ++j) {
            const uint64_t i = tile * kListTile + uint64_t(j) * kThreads + threadIdx.x;
            key[j] = i < total ? tk[i] : 0ull;
            occ |= (key[j] != 0 ? 1u : 0u) << j;
        }
        uint32_t unused;
        uint64_t o = offsets[tile] + block_exclusive_scan<kThreads>(uint32_t(__popc(occ)), s_scan, unused);
#pragma unroll
        for (int j = 0; j < kListIpt; ++j) {
            if ((occ >> j) & 1u) {
                list_keys[o] = key[j];
                list_slots[o] = uint32_t(tile * kListTile + uint64_t(j) * kThreads + threadIdx.x);
                ++o;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Distinct-count estimate of the leaf masks (HyperLogLog: Flajolet, Fusy,
// Gandouet, Meunier 2007), so the leaf table is sized by unique leaves rather
// than by leaf positions and its slot arrays stay cache-sized.  Only the table
// size depends on it; the result does not.
// ---------------------------------------------------------------------------
constexpr int kSketchIpt = 8;

__global__ __launch_bounds__(kThreads) void k_sketch(const uint64_t* __restrict__ values,
                                                     const uint64_t* __restrict__ count_ptr,
                                                     uint32_t* __restrict__ regs)
{
    __shared__ uint32_t s_reg[kSketchRegs];
    for (uint32_t i = threadIdx.x; i < uint32_t(kSketchRegs); i += kThreads) s_reg[i] = 0;
    __syncthreads();
    const uint64_t n = *count_ptr;
    // kSketchIpt values per thread and round, all loaded before any is used
    // (one workgroup per CU holds too few waves to hide one load at a time).
    // Positions past the end read the last value again: a register maximum
    // does not change when a value is counted twice.
    constexpr uint64_t kRound = uint64_t(kThreads) * kSketchIpt;
    for (uint64_t b = uint64_t(blockIdx.x) * kRound; b < n; b += uint64_t(gridDim.x) * kRound) {
        uint64_t v[kSketchIpt];
#pragma unroll
        for (int i = 0; i < kSketchIpt; ++i)
            v[i] = __builtin_nontemporal_load(values + min(b + uint64_t(i) * kThreads + threadIdx.x, n - 1));
#pragma unroll
        for (int i = 0; i < kSketchIpt; ++i) {
            // A mixer unrelated to the table's home-slot hash.
            const uint64_t h = fmix64(v[i] * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull);
            const uint32_t idx = uint32_t(h >> (64 - kSketchBits));
            const uint64_t rest = h << kSketchBits;
            const uint32_t rho = rest ? uint32_t(__clzll((long long)rest)) + 1u : uint32_t(64 - kSketchBits + 1);
            atomicMax(&s_reg[idx], rho);
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < uint32_t(kSketchRegs); i += kThreads)
        if (s_reg[i]) atomicMax(&regs[i], s_reg[i]);
}

constexpr int kEstimateThreads = 1024;

__global__ __launch_bounds__(kEstimateThreads) void k_sketch_estimate(const uint32_t* __restrict__ regs,
                                                                      const uint64_t* __restrict__ bound_ptr,
                                                                      uint64_t* __restrict__ estimate)
{
    __shared__ double s_sum[kEstimateThreads / kWave];
    __shared__ uint32_t s_zero[kEstimateThreads / kWave];
    double sum = 0.0;
    uint32_t zeros = 0;
    for (uint32_t i = threadIdx.x; i < uint32_t(kSketchRegs); i += kEstimateThreads) {
        sum += ldexp(1.0, -int(regs[i]));
        zeros += regs[i] == 0 ? 1u : 0u;
    }
#pragma unroll
    for (int d = kWave / 2; d >= 1; d >>= 1) {
        sum += __shfl_xor(sum, d, kWave);
        zeros += __shfl_xor(zeros, d, kWave);
    }
    const uint32_t w = threadIdx.x / kWave;
    if (lane_id() == 0) {
        s_sum[w] = sum;
        s_zero[w] = zeros;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double total = 0.0;
        uint32_t z = 0;
        for (int i = 0; i < kEstimateThreads / kWave; ++i) {
            total += s_sum[i];
            z += s_zero[i];
        }
        const double m = double(kSketchRegs);
        double e = 0.7213 / (1.0 + 1.079 / m) * m * m / total;
        if (e <= 2.5 * m && z) e = m * log(m / double(z));  // small-range (linear counting) correction
        const uint64_t want = uint64_t(e * 1.25) + 16;
        const uint64_t bound = *bound_ptr;
        *estimate = want < bound ? want : bound;
    }
}

__global__ void k_rank_of_slot(const uint32_t* __restrict__ sorted_slots, const uint64_t* __restrict__ count_ptr,
                               uint32_t* __restrict__ rank_of_slot)
{
    const uint64_t n = *count_ptr;
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t r = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; r < n; r += stride)
        rank_of_slot[sorted_slots[r]] = uint32_t(r);
}

// Ordered-table probe: home = top lg bits of the hash, linear probing forward
// into an overflow tail (no wrap), so every key sits at or after its home slot.
// `won` tells whether this call placed the key.
__device__ __forceinline__ uint32_t ordered_probe(uint64_t* __restrict__ tk, uint32_t lg, uint64_t limit, uint64_t h,
                                                  bool& won, unsigned long long* err)
{
    won = false;
    uint64_t slot = h >> (64 - lg);
    for (; slot < limit; ++slot) {
        const uint64_t cur = probe_load(&tk[slot]);
        if (cur == h) return uint32_t(slot);
        if (cur == 0) {
            const uint64_t prev = atomicCAS((unsigned long long*)&tk[slot], 0ull, (unsigned long long)h);
            if (prev == 0) {
                won = true;
                return uint32_t(slot);
            }
            if (prev == h) return uint32_t(slot);
        }
    }
    atomicOr(err, (unsigned long long)kErrTableOverflow);
    return uint32_t(limit - 1);
}

// Ordered-table insert; the placing call records p as the key's representative.
__device__ __forceinline__ uint32_t ordered_insert(uint64_t* __restrict__ tk, uint32_t* __restrict__ rep, uint32_t lg,
                                                   uint64_t limit, uint64_t h, uint32_t p, BuildCounters* c)
{
    bool won;
    const uint32_t slot = ordered_probe(tk, lg, limit, h, won, (unsigned long long*)&c->error);
    if (won) rep[slot] = p;
    return slot;
}

// ---------------------------------------------------------------------------
// Level build: children (Morton-sorted codes of the level below, their table
// slots) -> parents.  A parent belongs to the thread holding its first child and
// has at most 8 children, so a run extends at most 7 keys past the thread.
// ---------------------------------------------------------------------------
constexpr int kBuildIpt = 4;
constexpr int kBuildTile = kThreads * kBuildIpt;

__device__ __forceinline__ void parent_runs(const uint64_t* __restrict__ codes, uint64_t n, uint64_t first,
                                            uint64_t (&code)[kBuildIpt], uint32_t& starts, uint32_t& gcount)
{
    // Unconditional loads (positions past the end read the last code and are
    // never starts), so that all of them are in flight at once.  n >= 1 here.
#pragma unroll
    for (int j = 0; j < kBuildIpt; ++j) code[j] = codes[min(first + j, n - 1)];
    const uint64_t before = first > 0 ? codes[min(first - 1, n - 1)] : ~0ull;
    starts = 0;
    gcount = 0;
#pragma unroll
    for (int j = 0; j < kBuildIpt; ++j) {
        const uint64_t prev = j ? code[j - 1] : before;
        const bool head = first + j == 0;
        if (first + j < n && (head || (code[j] >> 3) != (prev >> 3))) {
            starts |= 1u << j;
            gcount += (head || (code[j] >> 6) != (prev >> 6)) ? 1u : 0u;
        }
    }
}

// With clear.keys set, the kernel also clears the level's table (as
// k_table_clear: the table's node count is not the one this kernel adds to).
__global__ __launch_bounds__(kThreads) void k_level_count(const uint64_t* __restrict__ codes,
                                                          const uint64_t* __restrict__ n_ptr,
                                                          uint32_t* __restrict__ counts,
                                                          uint64_t* __restrict__ next_count, TableClearArgs clear)
{
    if (clear.keys) {
        const uint64_t cnt = *clear.count;
        const uint32_t need = ceil_log2_u64((cnt ? cnt : 1ull) << clear.load_shift);
        const uint32_t lg = need > 8u ? need : 8u;
        const uint64_t total = (1ull << lg) + clear.overflow;
        if (blockIdx.x == 0 && threadIdx.x == 0) *clear.log2_out = lg;
        const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
        for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += stride) clear.keys[i] = 0;
    }
    __shared__ uint64_t s_red[kThreads / kWave];
    const uint64_t n = *n_ptr;
    const uint64_t ntiles = (n + kBuildTile - 1) / kBuildTile;
    uint64_t gtot = 0;
    for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        // The same run starts as parent_runs, counted with coalesced reads:
        // lane i of a round reads code i and its predecessor (all loads
        // issued first, indices clamped into [0, n)).
        const uint64_t base = tile * kBuildTile + threadIdx.x;
        uint64_t cv[kBuildIpt], pv[kBuildIpt];
#pragma unroll
        for (int j = 0; j < kBuildIpt; ++j) {
            const uint64_t i = base + uint64_t(j) * kThreads;
            cv[j] = codes[min(i, n - 1)];
            pv[j] = codes[i > 0 ? min(i - 1, n - 1) : 0];
        }
        uint32_t starts = 0, gcount = 0;
#pragma unroll
        for (int j = 0; j < kBuildIpt; ++j) {
            const uint64_t i = base + uint64_t(j) * kThreads;
            if (i < n && (i == 0 || (cv[j] >> 3) != (pv[j] >> 3))) {
                ++starts;
                gcount += (i == 0 || (cv[j] >> 6) != (pv[j] >> 6)) ? 1u : 0u;
            }
        }
        const uint64_t cnt = block_reduce_sum<kThreads>(starts, s_red);
        gtot += gcount;
        if (threadIdx.x == 0) counts[tile] = uint32_t(cnt);
    }
    if (next_count) {
        gtot = block_reduce_sum<kThreads>(gtot, s_red);
        if (threadIdx.x == 0 && gtot) atomicAdd((unsigned long long*)next_count, (unsigned long long)gtot);
    }
}

struct ParentAcc {
    uint64_t code;
    uint32_t mask;
    uint32_t child[8];
    uint64_t hash[8];
};

__device__ __forceinline__ void acc_reset(ParentAcc& a, uint64_t code)
{
    a.code = code;
    a.mask = 0;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        a.child[s] = kEmptyChild;
        a.hash[s] = 0;
    }
}

__device__ __forceinline__ void acc_add(ParentAcc& a, uint32_t slot, uint32_t rank, uint64_t hash)
{
    a.mask |= 1u << slot;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        a.child[s] = slot == uint32_t(s) ? rank : a.child[s];
        a.hash[s] = slot == uint32_t(s) ? hash : a.hash[s];
    }
}

// Writes a parent's record and returns its hash.
__device__ __forceinline__ uint64_t acc_emit(const ParentAcc& a, uint64_t p, const LevelBuildArgs& g)
{
    g.parent_codes[p] = a.code;
    g.parent_masks[p] = uint8_t(a.mask);
    uint4* dst = reinterpret_cast<uint4*>(g.parent_children + 8 * p);
    dst[0] = make_uint4(a.child[0], a.child[1], a.child[2], a.child[3]);
    dst[1] = make_uint4(a.child[4], a.child[5], a.child[6], a.child[7]);
    const uint64_t h = hash_children(a.hash);
    if (g.parent_hashes) g.parent_hashes[p] = h;
    return h;
}

// A thread owns kBuildIpt consecutive children; their loads (or rank and hash
// gathers) are issued together.  The tile's parent hashes are staged in LDS and
// inserted into the level's table one thread per parent.
__global__ __launch_bounds__(kThreads) void k_level_build(LevelBuildArgs g)
{
    __shared__ uint32_t s_scan[kThreads / kWave];
    __shared__ uint64_t s_hash[kBuildTile];
    __shared__ uint32_t s_elect[kThreads / kWave][kElectBuckets];
    const uint64_t n = *g.child_count;
    const uint64_t ntiles = (n + kBuildTile - 1) / kBuildTile;
    const uint32_t lg = *g.table_log2;
    const uint64_t limit = (1ull << lg) + kTableOverflow;
    const uint32_t* __restrict__ slots = g.child_slots;
    const uint32_t* __restrict__ rank_of_slot = g.child_rank_of_slot;
    const uint64_t* __restrict__ hashes = g.child_hashes;
    const uint64_t* __restrict__ seq = g.child_hash_seq;
    for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const uint64_t first = tile * kBuildTile + uint64_t(threadIdx.x) * kBuildIpt;
        uint64_t code[kBuildIpt];
        uint32_t starts, gcount;
        parent_runs(g.child_codes, n, first, code, starts, gcount);
        uint32_t tile_parents;
        const uint32_t local = block_exclusive_scan<kThreads>(__popc(starts), s_scan, tile_parents);
        const uint64_t p0 = g.offsets[tile];

        // Indices are clamped into [0, n): loads past the end read a valid child
        // that is then ignored.
        // rank[] holds the child pointers: slots with sequential hashes, else ranks.
        uint32_t rank[kBuildIpt];
        uint64_t hsh[kBuildIpt];
        if (seq) {
#pragma unroll
            for (int j = 0; j < kBuildIpt; ++j) {
                rank[j] = slots[min(first + j, n - 1)];
                hsh[j] = seq[min(first + j, n - 1)];
            }
        } else {
#pragma unroll
            for (int j = 0; j < kBuildIpt; ++j) rank[j] = rank_of_slot[slots[min(first + j, n - 1)]];
#pragma unroll
            for (int j = 0; j < kBuildIpt; ++j) hsh[j] = hashes[rank[j]];
        }

        uint32_t q = local;
        bool open = false;
        ParentAcc acc;
        acc_reset(acc, 0);
#pragma unroll
        for (int j = 0; j < kBuildIpt; ++j) {
            if ((starts >> j) & 1u) {
                if (open) {
                    s_hash[q] = acc_emit(acc, p0 + q, g);
                    ++q;
                }
                open = true;
                acc_reset(acc, code[j] >> 3);
            }
            if (open && first + j < n) acc_add(acc, uint32_t(code[j] & 7), rank[j], hsh[j]);
        }
        if (open) {
            // The last parent may take children past this thread's (at most 7).
            for (uint64_t i = first + kBuildIpt; i < n; ++i) {
                const uint64_t cc = g.child_codes[i];
                if ((cc >> 3) != acc.code) break;
                if (seq) {
                    acc_add(acc, uint32_t(cc & 7), slots[i], seq[i]);
                } else {
                    const uint32_t r = rank_of_slot[slots[i]];
                    acc_add(acc, uint32_t(cc & 7), r, hashes[r]);
                }
            }
            s_hash[q] = acc_emit(acc, p0 + q, g);
        }
        __syncthreads();
        // Equal parents (identical subtrees) often sit side by side: one lane
        // per distinct hash in a wave probes the table, the others take its slot.
        for (uint32_t k0 = 0; k0 < tile_parents; k0 += kThreads) {
            const uint32_t k = k0 + threadIdx.x;
            const bool valid = k < tile_parents;
            const uint64_t h = valid ? s_hash[k] : 0ull;
            bool follower;
            const uint32_t leader = elect_leader(h, valid, s_elect[threadIdx.x / kWave], follower);
            uint32_t slot = 0;
            if (valid && !follower)
                slot = ordered_insert(g.table_keys, g.table_rep, lg, limit, h, uint32_t(p0 + k), g.counters);
            slot = uint32_t(__shfl(int(slot), int(leader), kWave));
            if (valid) g.parent_slots[p0 + k] = slot;
        }
        __syncthreads();  // s_hash is rewritten by the next tile
    }
}

// ---------------------------------------------------------------------------
// Small top levels in one workgroup.  Per level (top down to 0): parents are
// the runs of equal code >> 3 among the children; one thread per parent reads
// its (at most 8) children and forms its mask, child ranks and hash; the
// (hash, parent) pairs are sorted in LDS (bitonic), equal hashes get one rank,
// and the level is written in ascending hash order.  The parents' ranks are
// the next level's child ranks.
// ---------------------------------------------------------------------------
constexpr int kTailThreads = 1024;

__global__ __launch_bounds__(kTailThreads) void k_level_tail(TailArgs a)
{
    __shared__ uint64_t s_key[kTailMaxParents];
    __shared__ uint16_t s_idx[kTailMaxParents];
    __shared__ uint32_t s_start[kTailMaxParents + 1];
    __shared__ uint32_t s_scan[kTailThreads / kWave];
    __shared__ uint64_t s_red[kTailThreads / kWave];
    const uint32_t tid = threadIdx.x;
    BuildCounters* c = a.counters;
    uint32_t C = uint32_t(*a.child_count);
    int cur = 0;
    for (int l = a.top; l >= 0; --l, cur ^= 1) {
        const bool first = l == a.top;
        const uint64_t* ccodes = first ? a.child_codes : a.codes[cur ^ 1];
        const uint32_t* cranks = first ? nullptr : a.ranks[cur ^ 1];
        const uint64_t* chashes = first ? a.child_hashes : a.out_hashes[l + 1];
        // Child rank and hash of child i.
        auto child_rank = [&](uint32_t i) -> uint32_t {
            return first ? a.child_rank_of_slot[a.child_slots[i]] : cranks[i];
        };

        // 1. Run heads (parents) in child order.
        uint32_t M = 0;
        for (uint32_t base = 0; base < C; base += kTailThreads) {
            const uint32_t i = base + tid;
            bool head = false;
            if (i < C) head = i == 0 || (ccodes[i] >> 3) != (ccodes[i - 1] >> 3);
            uint32_t total;
            const uint32_t ex = block_exclusive_scan<kTailThreads>(head ? 1u : 0u, s_scan, total);
            if (head) s_start[M + ex] = i;
            M += total;
        }
        if (tid == 0) s_start[M] = C;
        __syncthreads();

        // 2. Children scatter their rank and hash into their parent's slot (the
        //    gathers of all children are independent), then one thread per
        //    parent reads its 8 slots and forms mask, record and hash.
        uint64_t* pcodes = a.codes[cur];
        for (uint32_t k = tid; k < 8 * M; k += kTailThreads) {
            a.children[k] = kEmptyChild;
            a.child_hash[k] = 0;
        }
        __syncthreads();
#pragma unroll 4
        for (uint32_t i = tid; i < C; i += kTailThreads) {
            // Parent of child i: the last run start <= i.
            uint32_t lo = 0, hi = M;
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) / 2;
                if (s_start[mid] <= i) lo = mid;
                else hi = mid;
            }
            const uint32_t slot = uint32_t(ccodes[i] & 7);
            const uint32_t r = child_rank(i);
            a.children[8 * lo + slot] = r;
            a.child_hash[8 * lo + slot] = chashes[r];
        }
        __syncthreads();
        for (uint32_t p = tid; p < M; p += kTailThreads) {
            const uint4* cs = reinterpret_cast<const uint4*>(a.children + 8 * uint64_t(p));
            const uint4 c0 = cs[0], c1 = cs[1];
            const uint32_t ch[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
            uint32_t mask = 0;
            uint64_t hs[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                mask |= ch[k] != kEmptyChild ? 1u << k : 0u;
                hs[k] = a.child_hash[8 * uint64_t(p) + k];
            }
            pcodes[p] = ccodes[s_start[p]] >> 3;
            a.masks[p] = uint8_t(mask);
            s_key[p] = hash_children(hs);
            s_idx[p] = uint16_t(p);
        }
        // Pad to a power of two; (max key, index 0xFFFF) sorts after every parent.
        uint32_t n2 = 1;
        while (n2 < M) n2 <<= 1;
        for (uint32_t p = M + tid; p < n2; p += kTailThreads) {
            s_key[p] = ~0ull;
            s_idx[p] = 0xFFFF;
        }
        __syncthreads();

        // 3. Bitonic sort of (key, index), ascending.
        for (uint32_t size = 2; size <= n2; size <<= 1) {
            for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
                for (uint32_t t = tid; t < n2 / 2; t += kTailThreads) {
                    const uint32_t lo = 2 * t - (t & (stride - 1));
                    const uint32_t hi = lo + stride;
                    const bool up = (lo & size) == 0;
                    const uint64_t ka = s_key[lo], kb = s_key[hi];
                    const uint16_t ia = s_idx[lo], ib = s_idx[hi];
                    const bool greater = ka > kb || (ka == kb && ia > ib);
                    if (greater == up) {
                        s_key[lo] = kb;
                        s_key[hi] = ka;
                        s_idx[lo] = ib;
                        s_idx[hi] = ia;
                    }
                }
                __syncthreads();
            }
        }

        // 4. Unique ranks; the first of each equal-hash run represents it.
        uint32_t* pranks = a.ranks[cur];
        uint32_t Q = 0;
        uint64_t words = 0, bad = 0;
        for (uint32_t base = 0; base < M; base += kTailThreads) {
            const uint32_t q = base + tid;
            bool lead = false;
            if (q < M) lead = q == 0 || s_key[q] != s_key[q - 1];
            uint32_t total;
            const uint32_t ex = block_exclusive_scan<kTailThreads>(lead ? 1u : 0u, s_scan, total);
            if (q < M) {
                const uint32_t r = Q + ex + (lead ? 1u : 0u) - 1u;
                const uint32_t p = s_idx[q];
                pranks[p] = r;
                if (lead) {
                    const uint32_t m = a.masks[p];
                    a.out_hashes[l][r] = s_key[q];
                    a.out_masks[l][r] = uint8_t(m);
                    const uint4* src = reinterpret_cast<const uint4*>(a.children + 8 * uint64_t(p));
                    uint4* dst = reinterpret_cast<uint4*>(a.out_children[l] + 8 * uint64_t(r));
                    dst[0] = src[0];
                    dst[1] = src[1];
                    words += 1 + __popc(m);
                } else if (a.check_collisions) {
                    // Compare with the run's representative (content must be equal).
                    uint32_t f = q;
                    while (f > 0 && s_key[f - 1] == s_key[q]) --f;
                    const uint32_t rp = s_idx[f];
                    bool diff = a.masks[p] != a.masks[rp];
                    for (int k = 0; k < 8; ++k) diff |= a.children[8 * p + k] != a.children[8 * rp + k];
                    bad += diff ? 1 : 0;
                }
            }
            Q += total;
        }
        words = block_reduce_sum<kTailThreads>(words, s_red);
        bad = block_reduce_sum<kTailThreads>(bad, s_red);
        if (tid == 0) {
            c->pre[l] = M;
            c->uniq[l] = Q;
            c->words[l] = words;
            if (bad) c->collisions += bad;
        }
        // The next level reads this level's codes, ranks and hashes (the
        // barrier orders the global writes within the workgroup).
        __syncthreads();
        C = M;
    }
}

// Exact-content check of every parent against its table representative.
__global__ void k_level_check(const uint32_t* __restrict__ parent_slots, const uint32_t* __restrict__ table_rep,
                              const uint8_t* __restrict__ masks, const uint32_t* __restrict__ children,
                              const uint64_t* __restrict__ count_ptr, uint64_t* __restrict__ collisions)
{
    const uint64_t n = *count_ptr;
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    uint32_t bad = 0;
    for (uint64_t p = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; p < n; p += stride) {
        const uint32_t r = table_rep[parent_slots[p]];
        bool diff = masks[p] != masks[r];
        for (int s = 0; s < 8; ++s) diff |= children[8 * p + s] != children[8 * uint64_t(r) + s];
        bad += diff ? 1u : 0u;
    }
    if (bad) atomicAdd((unsigned long long*)collisions, (unsigned long long)bad);
}

// ---------------------------------------------------------------------------
// Level extract: ordered table -> the level's arrays in ascending hash order.
// For occupied slot k in probe cluster [js, je]:
//   rank(k) = (#occupied slots before k) - (k - js) + #{keys in cluster < key(k)}
// Slots are read through an LDS window with a halo; clusters reaching past it
// fall back to global loads.
// ---------------------------------------------------------------------------
constexpr int kExtractIpt = 1;
constexpr int kExtractTile = kThreads * kExtractIpt;
constexpr int kHalo = 128;

__global__ __launch_bounds__(kThreads) void k_extract_count(const uint64_t* __restrict__ table_keys,
                                                            const uint32_t* __restrict__ table_log2,
                                                            uint32_t* __restrict__ counts,
                                                            uint64_t* __restrict__ slots_out)
{
    __shared__ uint64_t s_red[kThreads / kWave];
    const uint64_t total = (1ull << *table_log2) + kTableOverflow;
    const uint64_t ntiles = (total + kExtractTile - 1) / kExtractTile;
    if (blockIdx.x == 0 && threadIdx.x == 0) *slots_out = total;
    for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const uint64_t first = tile * kExtractTile;
        uint32_t occ = 0;
        // Unconditional loads (slots past the end read the last slot and are
        // not counted), all in flight at once.
        uint64_t v[kExtractIpt];
#pragma unroll
        for (int k = 0; k < kExtractIpt; ++k) v[k] = table_keys[min(first + uint64_t(k) * kThreads + threadIdx.x, total - 1)];
#pragma unroll
        for (int k = 0; k < kExtractIpt; ++k)
            occ += (first + uint64_t(k) * kThreads + threadIdx.x < total && v[k] != 0) ? 1u : 0u;
        const uint64_t cnt = block_reduce_sum<kThreads>(occ, s_red);
        if (threadIdx.x == 0) counts[tile] = uint32_t(cnt);
    }
}

__global__ __launch_bounds__(kThreads) void k_level_extract(LevelExtractArgs g)
{
    __shared__ uint64_t s_win[kExtractTile + 2 * kHalo];
    __shared__ uint32_t s_scan[kThreads / kWave];
    __shared__ uint64_t s_red[kThreads / kWave];
    const uint32_t lg = *g.table_log2;
    const uint64_t total = (1ull << lg) + kTableOverflow;
    const uint64_t ntiles = (total + kExtractTile - 1) / kExtractTile;
    uint64_t wsum = 0;
    for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int64_t tile_base = int64_t(tile) * kExtractTile;
        const int64_t win_lo = tile_base - kHalo;
        const int64_t win_hi = tile_base + kExtractTile + kHalo;
        // Unconditional loads of the window (clamped into the table; slots
        // outside it read as empty), all in flight at once.
        constexpr int kWinPerThread = (kExtractTile + 2 * kHalo + kThreads - 1) / kThreads;
        uint64_t wv[kWinPerThread];
#pragma unroll
        for (int q = 0; q < kWinPerThread; ++q) {
            const int64_t s = win_lo + q * kThreads + int64_t(threadIdx.x);
            wv[q] = g.table_keys[min(max(s, int64_t(0)), int64_t(total) - 1)];
        }
#pragma unroll
        for (int q = 0; q < kWinPerThread; ++q) {
            const int i = q * kThreads + int(threadIdx.x);
            const int64_t s = win_lo + i;
            if (i < kExtractTile + 2 * kHalo) s_win[i] = (s >= 0 && s < int64_t(total)) ? wv[q] : 0;
        }
        __syncthreads();
        auto key_at = [&](int64_t s) -> uint64_t {
            if (s < 0 || s >= int64_t(total)) return 0;
            if (s >= win_lo && s < win_hi) return s_win[s - win_lo];
            return g.table_keys[s];
        };
        const int64_t first = tile_base + int64_t(threadIdx.x) * kExtractIpt;
        uint32_t occ = 0;
#pragma unroll
        for (int j = 0; j < kExtractIpt; ++j) occ |= (s_win[first + j - win_lo] != 0 ? 1u : 0u) << j;
        uint32_t unused;
        const uint32_t local = block_exclusive_scan<kThreads>(__popc(occ), s_scan, unused);
        uint64_t before = g.offsets[tile] + local;  // occupied slots before the current one
        // Representatives' records, gathered up front (unoccupied slots read
        // parent 0, which is then ignored).
        uint32_t mask[kExtractIpt];
        uint4 ch0[kExtractIpt], ch1[kExtractIpt];
#pragma unroll
        for (int j = 0; j < kExtractIpt; ++j) {
            const uint32_t r = g.table_rep[min(uint64_t(first + j), total - 1)];
            const uint64_t rep = ((occ >> j) & 1u) ? r : 0u;
            mask[j] = g.parent_masks[rep];
            const uint4* src = reinterpret_cast<const uint4*>(g.parent_children + 8 * rep);
            ch0[j] = src[0];
            ch1[j] = src[1];
        }
        if (g.child_rank_of_slot) {
            // Child slots of the level below -> its ranks.  Only occupied
            // slots' real children are translated (the others hold parent 0's
            // record, which may be unset); every gather is issued, reading
            // rank 0 for the rest, so that all are in flight at once.
            const uint32_t* __restrict__ rk = g.child_rank_of_slot;
            uint32_t r[kExtractIpt][8];
#pragma unroll
            for (int j = 0; j < kExtractIpt; ++j) {
                const bool live = (occ >> j) & 1u;
                const uint32_t c[8] = {ch0[j].x, ch0[j].y, ch0[j].z, ch0[j].w, ch1[j].x, ch1[j].y, ch1[j].z, ch1[j].w};
#pragma unroll
                for (int e = 0; e < 8; ++e) r[j][e] = rk[live && c[e] != kEmptyChild ? c[e] : 0u];
            }
            auto tr = [](uint32_t v, uint32_t rank) { return v == kEmptyChild ? v : rank; };
#pragma unroll
            for (int j = 0; j < kExtractIpt; ++j) {
                ch0[j] = make_uint4(tr(ch0[j].x, r[j][0]), tr(ch0[j].y, r[j][1]), tr(ch0[j].z, r[j][2]),
                                    tr(ch0[j].w, r[j][3]));
                ch1[j] = make_uint4(tr(ch1[j].x, r[j][4]), tr(ch1[j].y, r[j][5]), tr(ch1[j].z, r[j][6]),
                                    tr(ch1[j].w, r[j][7]));
            }
        }
#pragma unroll
        for (int j = 0; j < kExtractIpt; ++j) {
            if ((occ >> j) & 1u) {
                const int64_t k = first + j;
                const uint64_t key = s_win[k - win_lo];
                int64_t js = k;
                while (key_at(js - 1) != 0) --js;
                int64_t je = k;
                while (key_at(je + 1) != 0) ++je;
                uint64_t lt = 0;
                for (int64_t x = js; x <= je; ++x) lt += key_at(x) < key ? 1 : 0;
                const uint64_t rank = before - uint64_t(k - js) + lt;
                g.rank_of_slot[k] = uint32_t(rank);
                g.out_hashes[rank] = key;
                g.out_masks[rank] = uint8_t(mask[j]);
                uint4* dst = reinterpret_cast<uint4*>(g.out_children + 8 * rank);
                dst[0] = ch0[j];
                dst[1] = ch1[j];
                wsum += 1 + __popc(mask[j]);
                ++before;
            }
        }
        __syncthreads();  // the window is rewritten by the next tile
    }
    const uint64_t wtot = block_reduce_sum<kThreads>(wsum, s_red);
    if (threadIdx.x == 0 && wtot) atomicAdd((unsigned long long*)g.w_count, (unsigned long long)wtot);
}

// ---------------------------------------------------------------------------
// Final layout.  Node g of the level concatenation (levels 0.., in order) starts
// at word positions[g] = exclusive prefix of (1 + popc(mask)); leaves follow.
// ---------------------------------------------------------------------------
constexpr int kFinalIpt = 16;
constexpr int kFinalTile = kThreads * kFinalIpt;
static_assert(kBuildTile == kBuildTileItems && kExtractTile == kExtractTileItems &&
                  kFinalTile == kFinalTileItems && kListTile == kListTileItems,
              "tile sizes shared with the scratch planner");

// Per-level starts and arrays of the layout, staged in LDS once per block:
// a lane's level lookup then reads LDS rather than indexing the kernel
// argument block by a per-lane level (a dependent memory load per step).
struct FinalLds {
    uint64_t first[SVDAG_MAX_LEVELS + 1];  // first node of each level; [num_levels] = total_nodes
    const uint8_t* masks[SVDAG_MAX_LEVELS];
    const uint32_t* children[SVDAG_MAX_LEVELS];
};

// Every thread of the block must call it (one barrier).
__device__ __forceinline__ void stage_levels(const FinalArgs& a, FinalLds& s)
{
    const int t = int(threadIdx.x);
    if (t < a.num_levels) {
        s.first[t] = a.levels[t].first_node;
        s.masks[t] = a.levels[t].masks;
        s.children[t] = a.levels[t].children;
    }
    if (t == a.num_levels) s.first[t] = a.total_nodes;
    __syncthreads();
}

// Level of node g (< total_nodes): the last level starting at or before g;
// empty levels start where the next one does and are passed over.
__device__ __forceinline__ int level_of(const FinalLds& s, int num_levels, uint64_t g)
{
    int l = 0;
    while (l + 1 < num_levels && g >= s.first[l + 1]) ++l;
    return l;
}

// Word counts (1 + popc(mask)) of the thread's nodes.  The levels of the
// nodes are found first (they depend only on the index), so that all mask
// loads are in flight at once; nodes past the end read the last node's mask
// and count 0.
__device__ __forceinline__ uint32_t node_words(const FinalArgs& a, const FinalLds& s, uint64_t first,
                                               uint32_t (&sizes)[kFinalIpt])
{
    const uint64_t last = a.total_nodes - 1;  // total_nodes >= 1 (the root)
    int l = level_of(s, a.num_levels, min(first, last));
    const uint8_t* src[kFinalIpt];
#pragma unroll
    for (int j = 0; j < kFinalIpt; ++j) {
        const uint64_t gi = min(first + j, last);
        while (gi >= s.first[l + 1]) ++l;
        src[j] = s.masks[l] + (gi - s.first[l]);
    }
    uint8_t m[kFinalIpt];
#pragma unroll
    for (int j = 0; j < kFinalIpt; ++j) m[j] = *src[j];
    uint32_t sum = 0;
#pragma unroll
    for (int j = 0; j < kFinalIpt; ++j) {
        sizes[j] = first + j < a.total_nodes ? 1 + __popc(uint32_t(m[j])) : 0u;
        sum += sizes[j];
    }
    return sum;
}

__global__ __launch_bounds__(kThreads) void k_final_count(FinalArgs a)
{
    __shared__ uint64_t s_red[kThreads / kWave];
    __shared__ FinalLds s_lv;
    stage_levels(a, s_lv);
    const uint64_t ntiles = (a.total_nodes + kFinalTile - 1) / kFinalTile;
    for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        uint32_t sizes[kFinalIpt];
        const uint32_t sum = node_words(a, s_lv, tile * kFinalTile + uint64_t(threadIdx.x) * kFinalIpt, sizes);
        const uint64_t tot = block_reduce_sum<kThreads>(sum, s_red);
        if (threadIdx.x == 0) a.counts[tile] = uint32_t(tot);
    }
}

__global__ __launch_bounds__(kThreads) void k_final_positions(FinalArgs a)
{
    __shared__ uint32_t s_scan[kThreads / kWave];
    __shared__ FinalLds s_lv;
    stage_levels(a, s_lv);
    const uint64_t ntiles = (a.total_nodes + kFinalTile - 1) / kFinalTile;
    for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const uint64_t first = tile * kFinalTile + uint64_t(threadIdx.x) * kFinalIpt;
        uint32_t sizes[kFinalIpt];
        const uint32_t sum = node_words(a, s_lv, first, sizes);
        uint32_t unused;
        const uint32_t local = block_exclusive_scan<kThreads>(sum, s_scan, unused);
        uint64_t pos = a.offsets[tile] + local;
#pragma unroll
        for (int j = 0; j < kFinalIpt; ++j) {
            if (first + j < a.total_nodes) a.positions[first + j] = uint32_t(pos);
            pos += sizes[j];
        }
    }
}

__global__ __launch_bounds__(kThreads) void k_final_write(FinalArgs a)
{
    __shared__ FinalLds s_lv;
    stage_levels(a, s_lv);
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t gi = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; gi < a.total_nodes; gi += stride) {
        const int l = level_of(s_lv, a.num_levels, gi);
        const uint64_t i = gi - s_lv.first[l];
        const uint32_t mask = s_lv.masks[l][i];
        const uint32_t pos = a.positions[gi];
        a.words[pos] = mask;
        const bool last = l + 1 == a.num_levels;
        const uint4* src = reinterpret_cast<const uint4*>(s_lv.children[l] + 8 * i);
        const uint4 c0 = src[0], c1 = src[1];
        const uint32_t ch[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
        // Every child's target first (absent children read node 0's position
        // and are dropped), so that all gathers are in flight at once.
        uint32_t target[8];
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const uint32_t c = (mask >> s) & 1u ? ch[s] : 0u;
            target[s] = last ? uint32_t(a.leaves_start + 2ull * c) : a.positions[s_lv.first[l + 1] + c];
        }
        uint32_t off = 1;
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            if ((mask >> s) & 1u) {
                a.words[pos + off] = target[s];
                ++off;
            }
        }
    }
    for (uint64_t r = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; r < a.num_leaves; r += stride) {
        const uint64_t leaf = a.leaves[r];
        a.words[a.leaves_start + 2 * r] = uint32_t(leaf);
        a.words[a.leaves_start + 2 * r + 1] = uint32_t(leaf >> 32);
    }
}

// ---------------------------------------------------------------------------
// Merge insert (MergeDAGs, src/DAGCompression.cu:1109-1320): the nodes of one
// level of many sub-DAGs, concatenated, enter one shared dedup table.  A node's
// hash depends only on its content, so it is reused; its children are remapped
// from sub-DAG-local indices to the merged level below through that level's
// slot and rank arrays.  Leaves (ordered == 0) go into the hashed leaf table.
// ---------------------------------------------------------------------------
__global__ void k_merge_insert(const MergeSource* __restrict__ src, uint32_t nsrc, uint64_t total,
                               const uint32_t* __restrict__ child_slot_of,
                               const uint32_t* __restrict__ child_rank_of_slot, uint8_t* __restrict__ masks_out,
                               uint32_t* __restrict__ children_out, uint64_t* __restrict__ tk,
                               uint32_t* __restrict__ trep, const uint32_t* __restrict__ log2_ptr,
                               uint32_t* __restrict__ slot_of, BuildCounters* __restrict__ c, int ordered,
                               uint64_t* __restrict__ list_keys, uint32_t* __restrict__ list_slots)
{
    const uint32_t lg = *log2_ptr;
    const uint64_t limit = (1ull << lg) + kTableOverflow;
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    // Wave-uniform trip count: the leaf mode's list reservation is a wave ballot.
    for (uint64_t base = uint64_t(blockIdx.x) * blockDim.x; base < total; base += stride) {
        const uint64_t gi = base + threadIdx.x;
        const bool active = gi < total;
        uint32_t lo = 0, hi = nsrc;  // src[lo].first <= gi < src[hi].first
        while (active && hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (src[mid].first <= gi) lo = mid;
            else hi = mid;
        }
        const MergeSource& s = src[lo];
        const uint64_t i = gi - s.first;
        if (!ordered) {
            bool won = false;
            uint32_t slot = 0;
            uint64_t key = 0;
            if (active) {
                key = s.hashes[i];
                slot = hashed_insert(tk, lg, key, won, (unsigned long long*)&c->error);
                slot_of[gi] = slot;
            }
            // One list reservation per wave for the keys this wave placed.
            const uint64_t wins = __ballot(won);
            if (wins) {
                const uint32_t leader = uint32_t(__ffsll((long long)wins) - 1);
                uint64_t start = 0;
                if (lane_id() == leader)
                    start = atomicAdd((unsigned long long*)&c->unique_leaves, (unsigned long long)__popcll(wins));
                start = __shfl(start, leader, kWave);
                if (won) {
                    const uint64_t idx = start + uint32_t(__popcll(wins & lanemask_lt()));
                    list_keys[idx] = key;
                    list_slots[idx] = slot;
                }
            }
            continue;
        }
        if (!active) continue;
        const uint64_t key = s.hashes[i];
        const uint4* cs = reinterpret_cast<const uint4*>(s.children + 8 * i);
        const uint4 a = cs[0], b = cs[1];
        uint32_t ch[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (ch[k] != kEmptyChild) ch[k] = child_rank_of_slot[child_slot_of[s.child_first + ch[k]]];
        masks_out[gi] = s.masks[i];
        uint4* dst = reinterpret_cast<uint4*>(children_out + 8 * gi);
        dst[0] = make_uint4(ch[0], ch[1], ch[2], ch[3]);
        dst[1] = make_uint4(ch[4], ch[5], ch[6], ch[7]);
        slot_of[gi] = ordered_insert(tk, trep, lg, limit, key, uint32_t(gi), c);
    }
}

// ---------------------------------------------------------------------------
// Resident DAG (served mode): one dedup table per level that persists across
// insertions.  A resident node is identified by its table slot; its record
// (mask, children as slots of the level below) is stored at that slot by the
// call that placed its key, and rep[slot] = slot, so the level extract reads
// a resident table like a build's.  Wave-aggregated counts of placed keys
// track each table's occupancy.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void count_placed(bool won, unsigned long long* occupied)
{
    const uint64_t wins = __ballot(won);
    if (wins && lane_id() == uint32_t(__ffsll((long long)wins) - 1))
        atomicAdd(occupied, (unsigned long long)__popcll(wins));
}

// Leaves of a round (unique masks) into the resident leaf table.
__global__ void k_resident_leaves(const uint64_t* __restrict__ leaves, uint64_t count, uint64_t* __restrict__ tk,
                                  uint32_t lg, uint32_t* __restrict__ slot_of, unsigned long long* __restrict__ occupied,
                                  unsigned long long* __restrict__ err)
{
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    // Wave-uniform trip count (count_placed is a wave ballot).
    for (uint64_t base = uint64_t(blockIdx.x) * blockDim.x; base < count; base += stride) {
        const uint64_t i = base + threadIdx.x;
        bool won = false;
        if (i < count) slot_of[i] = hashed_insert(tk, lg, leaves[i], won, err);
        count_placed(won, occupied);
    }
}

// Nodes of one round level into the resident table of that level; children
// (round ranks of the level below) become resident slots through child_slot_of.
__global__ void k_resident_level(ResidentLevelArgs a)
{
    const uint64_t limit = (1ull << a.lg) + kTableOverflow;
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t base = uint64_t(blockIdx.x) * blockDim.x; base < a.count; base += stride) {
        const uint64_t i = base + threadIdx.x;
        bool won = false;
        if (i < a.count) {
            const uint32_t slot = ordered_probe(a.table_keys, a.lg, limit, a.hashes[i], won, a.err);
            if (won) {
                const uint4* cs = reinterpret_cast<const uint4*>(a.children + 8 * i);
                const uint4 c0 = cs[0], c1 = cs[1];
                uint32_t ch[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
#pragma unroll
                for (int k = 0; k < 8; ++k)
                    if (ch[k] != kEmptyChild) ch[k] = a.child_slot_of[ch[k]];
                uint4* dst = reinterpret_cast<uint4*>(a.rec_children + 8 * uint64_t(slot));
                dst[0] = make_uint4(ch[0], ch[1], ch[2], ch[3]);
                dst[1] = make_uint4(ch[4], ch[5], ch[6], ch[7]);
                a.rec_masks[slot] = a.masks[i];
                a.table_rep[slot] = slot;
            }
            a.slot_of[i] = slot;
        }
        count_placed(won, a.occupied);
    }
}

// One thread per chunk slot s: walks the round's top levels along s's child
// indices; a chunk present in the round gets its root's resident slot and hash.
// A chunk at a slot that already holds one sets kErrSlotFilled.
constexpr int kRootThreads = 256;

__global__ __launch_bounds__(kRootThreads) void k_resident_roots(ResidentRootArgs a)
{
    const uint32_t slots = 1u << (3 * a.top_levels);
    for (uint32_t s = threadIdx.x; s < slots; s += kRootThreads) {
        uint64_t node = 0;
        bool present = a.has_root;
        for (int l = 0; l < a.top_levels && present; ++l) {
            const uint32_t b = (s >> (3 * (a.top_levels - 1 - l))) & 7u;
            if (!((a.masks[l][node] >> b) & 1u)) present = false;
            else node = a.children[l][8 * node + b];
        }
        uint64_t hash = 0;
        if (present) {
            const uint32_t slot = a.root_slot_of[node];
            if (a.root_slot[s] != kEmptyChild) {
                atomicOr(a.err, (unsigned long long)kErrSlotFilled);
            } else {
                a.root_slot[s] = slot;
                hash = a.root_table_keys[slot];
            }
        }
        a.root_hash[s] = hash;
    }
}

// One workgroup checks the device chunk offsets of a batched build and sets
// its key count (0 when they are invalid, so that no kernel reads past them).
constexpr int kCountThreads = 1024;

__global__ __launch_bounds__(kCountThreads) void k_batch_count(const uint64_t* __restrict__ offsets,
                                                               uint64_t num_chunks, uint64_t n_bound,
                                                               uint64_t* __restrict__ n_out,
                                                               unsigned long long* __restrict__ err)
{
    __shared__ uint32_t s_bad;
    if (threadIdx.x == 0) s_bad = offsets[0] != 0 || offsets[num_chunks] > n_bound;
    __syncthreads();
    for (uint64_t c = threadIdx.x; c < num_chunks; c += kCountThreads)
        if (offsets[c + 1] < offsets[c]) s_bad = 1;
    __syncthreads();
    if (threadIdx.x == 0) {
        *n_out = s_bad ? 0 : offsets[num_chunks];
        if (s_bad) atomicOr(err, (unsigned long long)kErrOffsets);
    }
}

// Keys of several device arrays, back to back.  A workgroup copies a tile of
// kGatherTile destination keys: one binary search finds the tile's first
// source, then the sources overlapping the tile are copied in turn, each as a
// contiguous, coalesced range.
constexpr uint32_t kGatherTile = 4096;

__global__ __launch_bounds__(kThreads) void k_gather_keys(const GatherSource* __restrict__ src, uint32_t nsrc,
                                                          uint64_t total, uint64_t* __restrict__ dst)
{
    __shared__ uint32_t s_first;
    const uint64_t ntiles = (total + kGatherTile - 1) / kGatherTile;
    for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const uint64_t t0 = tile * kGatherTile;
        const uint64_t t1 = min(t0 + kGatherTile, total);
        if (threadIdx.x == 0) {
            uint32_t lo = 0, hi = nsrc;  // src[lo].start <= t0 < src[hi].start
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) / 2;
                if (src[mid].start <= t0) lo = mid;
                else hi = mid;
            }
            s_first = lo;
        }
        __syncthreads();
        uint32_t sidx = s_first;
        for (uint64_t pos = t0; pos < t1; ++sidx) {
            const GatherSource g = src[sidx];
            const uint64_t end = min(t1, sidx + 1 < nsrc ? src[sidx + 1].start : total);
            for (uint64_t i = pos + threadIdx.x; i < end; i += kThreads)
                dst[i] = __builtin_nontemporal_load(g.keys + (i - g.start));
            pos = end;
        }
        __syncthreads();  // s_first is rewritten by the next tile
    }
}

// The filled chunk slots in slot (Morton) order: codes, root slots, count.
__global__ __launch_bounds__(kRootThreads) void k_resident_root_list(const uint32_t* __restrict__ root_slot,
                                                                     uint32_t slots, uint64_t* __restrict__ codes,
                                                                     uint32_t* __restrict__ list_slots,
                                                                     uint64_t* __restrict__ count)
{
    __shared__ uint32_t s_scan[kRootThreads / kWave];
    uint32_t placed = 0;
    for (uint32_t base = 0; base < slots; base += kRootThreads) {
        const uint32_t s = base + threadIdx.x;
        const bool filled = s < slots && root_slot[s] != kEmptyChild;
        uint32_t total;
        const uint32_t ex = block_exclusive_scan<kRootThreads>(filled ? 1u : 0u, s_scan, total);
        if (filled) {
            codes[placed + ex] = s;
            list_slots[placed + ex] = root_slot[s];
        }
        placed += total;
    }
    if (threadIdx.x == 0) *count = placed;
}

}  // namespace

// ---------------------------------------------------------------------------
// Host launchers.
// ---------------------------------------------------------------------------
void launch_fill_u64(hipStream_t s, uint64_t* p, uint64_t count, uint64_t value, uint64_t set_at, uint64_t set_value)
{
    if (count == 0) return;
    TimedLaunch t(kKFill, s);
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_for((count + kThreads - 1) / kThreads, 8)), dim3(kThreads), 0, s, p, count,
                       value, set_at, set_value);
}

void launch_scan_counts(hipStream_t s, const uint32_t* counts, uint64_t* offsets, const uint64_t* items_dev,
                        uint64_t per_tile, uint64_t* total_out, uint64_t items_host)
{
    TimedLaunch t(kKScanCounts, s);
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(kScanThreads), 0, s, counts, offsets, items_dev, per_tile, total_out,
                       items_host);
}

template <typename K>
void launch_leaf_reduce(hipStream_t s, const K* sorted, BuildCounters* c, uint32_t group_bits, uint64_t* reg_code,
                        uint64_t* reg_mask, uint64_t* lp_code, uint64_t* lp_mask, uint64_t lp_capacity,
                        uint32_t* tile_counts, uint64_t* tile_offsets, uint64_t n_bound, int leaf_level,
                        uint64_t code_mask, const ChunkPrefix& buckets)
{
    const uint64_t chunks = (n_bound + kReduceChunk - 1) / kReduceChunk;
    const uint64_t units = (chunks + kReduceWaves - 1) / kReduceWaves;
    {
        TimedLaunch t(kKLeafReduce, s);
        bool bucketed = false;
        if constexpr (sizeof(K) == 4) {
            if (buckets.offsets) {
                bucketed = true;
                if (buckets.num_chunks <= kLdsBuckets)
                    hipLaunchKernelGGL((k_leaf_reduce<K, kBucketsLds>),
                                       dim3(resident_grid(k_leaf_reduce<K, kBucketsLds>, kThreads, units)),
                                       dim3(kThreads), 0, s, sorted, &c->n, code_mask, group_bits, reg_code, reg_mask,
                                       tile_counts, &c->pre[leaf_level],
                                       reinterpret_cast<unsigned long long*>(&c->error), buckets);
                else
                    hipLaunchKernelGGL((k_leaf_reduce<K, kBucketsGlobal>),
                                       dim3(resident_grid(k_leaf_reduce<K, kBucketsGlobal>, kThreads, units)),
                                       dim3(kThreads), 0, s, sorted, &c->n, code_mask, group_bits, reg_code, reg_mask,
                                       tile_counts, &c->pre[leaf_level],
                                       reinterpret_cast<unsigned long long*>(&c->error), buckets);
            }
        }
        if (!bucketed)
            hipLaunchKernelGGL((k_leaf_reduce<K, kBucketsNone>),
                               dim3(resident_grid(k_leaf_reduce<K, kBucketsNone>, kThreads, units)), dim3(kThreads), 0,
                               s, sorted, &c->n, code_mask, group_bits, reg_code, reg_mask, tile_counts,
                               &c->pre[leaf_level], reinterpret_cast<unsigned long long*>(&c->error), ChunkPrefix{});
    }
    launch_scan_counts(s, tile_counts, tile_offsets, &c->n, kReduceChunk, &c->leaf_positions);
    {
        TimedLaunch t(kKLeafCompact, s);
        hipLaunchKernelGGL(k_leaf_compact, dim3(grid_for(units, 8)), dim3(kThreads), 0, s, reg_code, reg_mask, tile_counts,
                           tile_offsets, &c->n, lp_code, lp_mask, lp_capacity,
                           reinterpret_cast<unsigned long long*>(&c->error));
    }
}
template void launch_leaf_reduce<uint64_t>(hipStream_t, const uint64_t*, BuildCounters*, uint32_t, uint64_t*,
                                           uint64_t*, uint64_t*, uint64_t*, uint64_t, uint32_t*, uint64_t*, uint64_t,
                                           int, uint64_t, const ChunkPrefix&);
template void launch_leaf_reduce<uint32_t>(hipStream_t, const uint32_t*, BuildCounters*, uint32_t, uint64_t*,
                                           uint64_t*, uint64_t*, uint64_t*, uint64_t, uint32_t*, uint64_t*, uint64_t,
                                           int, uint64_t, const ChunkPrefix&);

void launch_table_clear(hipStream_t s, uint64_t* keys, const uint64_t* count_ptr, uint32_t* log2_out,
                        uint32_t load_shift, uint32_t overflow, uint64_t slot_bound)
{
    TimedLaunch t(kKTableClear, s);
    hipLaunchKernelGGL(k_table_clear, dim3(grid_for((slot_bound + kThreads - 1) / kThreads, 8)), dim3(kThreads), 0, s,
                       keys, count_ptr, log2_out, load_shift, overflow);
}

void launch_leaf_insert(hipStream_t s, const uint64_t* lp_mask, const uint64_t* count_ptr, uint64_t* table_keys,
                        const uint32_t* log2_ptr, uint32_t* lp_slot, uint64_t n_bound, unsigned long long* err)
{
    TimedLaunch t(kKLeafInsert, s);
    hipLaunchKernelGGL(k_leaf_insert, dim3(grid_for((n_bound + kThreads * kInsertIpt - 1) / (kThreads * kInsertIpt), 8)),
                       dim3(kThreads), 0, s, lp_mask, count_ptr, table_keys, log2_ptr, lp_slot, err);
}

void launch_slot_list(hipStream_t s, const uint64_t* table_keys, const uint32_t* log2_ptr, uint64_t slot_bound,
                      uint32_t* tile_counts, uint64_t* tile_offsets, uint64_t* slots_scratch, uint64_t* list_keys,
                      uint32_t* list_slots, uint64_t* list_count)
{
    const int grid = grid_for((slot_bound + kListTile - 1) / kListTile, 8);
    {
        TimedLaunch t(kKSlotList, s);
        hipLaunchKernelGGL(k_slot_count, dim3(grid), dim3(kThreads), 0, s, table_keys, log2_ptr, tile_counts,
                           slots_scratch);
    }
    launch_scan_counts(s, tile_counts, tile_offsets, slots_scratch, kListTile, list_count);
    {
        TimedLaunch t(kKSlotList, s);
        hipLaunchKernelGGL(k_slot_list, dim3(grid), dim3(kThreads), 0, s, table_keys, log2_ptr, tile_offsets,
                           list_keys, list_slots);
    }
}

void launch_sketch(hipStream_t s, const uint64_t* values, const uint64_t* count_ptr, uint32_t* regs, uint64_t n_bound)
{
    TimedLaunch t(kKSketch, s);
    // Few workgroups: each merges its registers into the global ones once.
    const uint64_t rounds = (n_bound + uint64_t(kThreads) * kSketchIpt - 1) / (uint64_t(kThreads) * kSketchIpt);
    hipLaunchKernelGGL(k_sketch, dim3(grid_for(rounds, 1)), dim3(kThreads), 0, s, values, count_ptr, regs);
}

void launch_sketch_estimate(hipStream_t s, const uint32_t* regs, const uint64_t* bound_ptr, uint64_t* estimate)
{
    TimedLaunch t(kKSketch, s);
    hipLaunchKernelGGL(k_sketch_estimate, dim3(1), dim3(kEstimateThreads), 0, s, regs, bound_ptr, estimate);
}

void launch_rank_of_slot(hipStream_t s, const uint32_t* sorted_slots, const uint64_t* count_ptr,
                         uint32_t* rank_of_slot, uint64_t n_bound)
{
    TimedLaunch t(kKRankOfSlot, s);
    hipLaunchKernelGGL(k_rank_of_slot, dim3(grid_for((n_bound + kThreads - 1) / kThreads, 8)), dim3(kThreads), 0, s,
                       sorted_slots, count_ptr, rank_of_slot);
}

void launch_level_build(hipStream_t s, const LevelBuildArgs& a, uint64_t child_bound, const TableClearArgs* clear)
{
    const uint64_t tiles = (child_bound + kBuildTile - 1) / kBuildTile;
    {
        TimedLaunch t(kKLevelCount, s);
        const uint64_t work = clear ? std::max(tiles, (clear->slot_bound + kThreads - 1) / kThreads) : tiles;
        const TableClearArgs ca = clear ? *clear : TableClearArgs{};  // by value: *clear may not outlive the call
        hipLaunchKernelGGL(k_level_count, dim3(resident_grid(k_level_count, kThreads, work)), dim3(kThreads), 0, s,
                           a.child_codes, a.child_count, a.tile_counts, a.next_count, ca);
    }
    launch_scan_counts(s, a.tile_counts, const_cast<uint64_t*>(a.offsets), a.child_count, kBuildTile, nullptr);
    {
        TimedLaunch t(kKLevelBuild, s);
        hipLaunchKernelGGL(k_level_build, dim3(resident_grid(k_level_build, kThreads, tiles)), dim3(kThreads), 0, s, a);
    }
}

void launch_level_tail(hipStream_t s, const TailArgs& a)
{
    TimedLaunch t(kKLevelTail, s);
    hipLaunchKernelGGL(k_level_tail, dim3(1), dim3(kTailThreads), 0, s, a);
}

void launch_level_extract(hipStream_t s, const LevelExtractArgs& a, uint64_t slot_bound)
{
    const uint64_t tiles = (slot_bound + kExtractTile - 1) / kExtractTile;
    {
        TimedLaunch t(kKExtractCount, s);
        hipLaunchKernelGGL(k_extract_count, dim3(resident_grid(k_extract_count, kThreads, tiles)), dim3(kThreads), 0,
                           s, a.table_keys, a.table_log2, a.tile_counts, a.slots_scratch);
    }
    launch_scan_counts(s, a.tile_counts, const_cast<uint64_t*>(a.offsets), a.slots_scratch, kExtractTile, a.q_count);
    {
        TimedLaunch t(kKLevelExtract, s);
        hipLaunchKernelGGL(k_level_extract, dim3(resident_grid(k_level_extract, kThreads, tiles)), dim3(kThreads), 0,
                           s, a);
    }
}

void launch_level_check(hipStream_t s, const uint32_t* parent_slots, const uint32_t* table_rep,
                        const uint8_t* parent_masks, const uint32_t* parent_children, const uint64_t* count_ptr,
                        uint64_t* collisions, uint64_t n_bound)
{
    TimedLaunch t(kKLevelCheck, s);
    hipLaunchKernelGGL(k_level_check, dim3(grid_for((n_bound + kThreads - 1) / kThreads, 8)), dim3(kThreads), 0, s,
                       parent_slots, table_rep, parent_masks, parent_children, count_ptr, collisions);
}

// ---------------------------------------------------------------------------
// Point query against the flat layout, walking it as the reference's tracer
// does (src/DAGTracer.cu:85-176): a node is its child mask word followed by
// one absolute word index per set slot; after depth - 2 nodes the index
// names a leaf, two words of 4x4x4 occupancy (low word first).  The slot of
// level l is bits [3(depth-1-l), 3(depth-l)) of the code.
// out[i]: 1 set, 0 not set (or the code is outside 2^(3*depth)), 2 the walk
// left the words (not a valid layout).
// ---------------------------------------------------------------------------
__global__ void k_query_voxels(const uint32_t* __restrict__ words, uint64_t n_words, int depth,
                               const uint64_t* __restrict__ codes, uint64_t n, uint8_t* __restrict__ out)
{
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint64_t c = codes[i];
        uint8_t r = 3;  // walk in progress
        if (n_words == 0 || (c >> (3 * depth)) != 0) r = 0;
        uint64_t idx = 0;
        for (int l = 0; r == 3 && l < depth - 2; ++l) {
            if (idx >= n_words) {
                r = 2;
                break;
            }
            const uint32_t mask = words[idx];
            const uint32_t slot = uint32_t(c >> (3 * (depth - 1 - l))) & 7u;
            if (mask > 0xFFu) {
                r = 2;
            } else if (!((mask >> slot) & 1u)) {
                r = 0;
            } else {
                const uint64_t at = idx + 1 + uint32_t(__popc(mask & ((1u << slot) - 1u)));
                if (at >= n_words) r = 2;
                else idx = words[at];
            }
        }
        if (r == 3) {
            if (idx + 1 >= n_words) {
                r = 2;
            } else {
                const uint32_t bit = uint32_t(c & 63u);
                r = uint8_t((words[idx + (bit >> 5)] >> (bit & 31u)) & 1u);
            }
        }
        out[i] = r;
    }
}

void launch_query_voxels(hipStream_t s, const uint32_t* words, uint64_t n_words, int depth, const uint64_t* codes,
                         uint64_t n, uint8_t* out)
{
    if (n == 0) return;
    TimedLaunch t(kKQueryVoxels, s);
    hipLaunchKernelGGL(k_query_voxels, dim3(grid_for((n + kThreads - 1) / kThreads, 8)), dim3(kThreads), 0, s, words,
                       n_words, depth, codes, n, out);
}

void launch_final(hipStream_t s, const FinalArgs& a)
{
    if (a.total_nodes == 0 && a.num_leaves == 0) return;
    if (a.total_nodes) {
        const uint64_t tiles = (a.total_nodes + kFinalTile - 1) / kFinalTile;
        {
            TimedLaunch t(kKFinalCount, s);
            hipLaunchKernelGGL(k_final_count, dim3(resident_grid(k_final_count, kThreads, tiles)), dim3(kThreads), 0, s,
                               a);
        }
        launch_scan_counts(s, a.counts, a.offsets, nullptr, kFinalTile, nullptr, a.total_nodes);
        {
            TimedLaunch t(kKFinalPositions, s);
            hipLaunchKernelGGL(k_final_positions, dim3(resident_grid(k_final_positions, kThreads, tiles)),
                               dim3(kThreads), 0, s, a);
        }
    }
    const uint64_t work = std::max(a.total_nodes, a.num_leaves);
    {
        TimedLaunch t(kKFinalWrite, s);
        hipLaunchKernelGGL(k_final_write, dim3(grid_for((work + kThreads - 1) / kThreads, 8)), dim3(kThreads), 0, s, a);
    }
}

void launch_resident_leaves(hipStream_t s, const uint64_t* leaves, uint64_t count, uint64_t* table_keys, uint32_t lg,
                            uint32_t* slot_of, unsigned long long* occupied, unsigned long long* err)
{
    if (count == 0) return;
    TimedLaunch t(kKResidentInsert, s);
    hipLaunchKernelGGL(k_resident_leaves, dim3(grid_for((count + kThreads - 1) / kThreads, 8)), dim3(kThreads), 0, s,
                       leaves, count, table_keys, lg, slot_of, occupied, err);
}

void launch_resident_level(hipStream_t s, const ResidentLevelArgs& a)
{
    if (a.count == 0) return;
    TimedLaunch t(kKResidentInsert, s);
    hipLaunchKernelGGL(k_resident_level, dim3(grid_for((a.count + kThreads - 1) / kThreads, 8)), dim3(kThreads), 0, s,
                       a);
}

void launch_batch_count(hipStream_t s, const uint64_t* offsets, uint64_t num_chunks, uint64_t n_bound, uint64_t* n_out,
                        unsigned long long* err)
{
    TimedLaunch t(kKFill, s);
    hipLaunchKernelGGL(k_batch_count, dim3(1), dim3(kCountThreads), 0, s, offsets, num_chunks, n_bound, n_out, err);
}

void launch_gather_keys(hipStream_t s, const GatherSource* src, uint32_t nsrc, uint64_t total, uint64_t* dst)
{
    if (total == 0 || nsrc == 0) return;
    TimedLaunch t(kKFill, s);
    hipLaunchKernelGGL(k_gather_keys, dim3(grid_for((total + kGatherTile - 1) / kGatherTile, 8)), dim3(kThreads), 0, s,
                       src, nsrc, total, dst);
}

void launch_resident_roots(hipStream_t s, const ResidentRootArgs& a)
{
    TimedLaunch t(kKResidentInsert, s);
    hipLaunchKernelGGL(k_resident_roots, dim3(1), dim3(kRootThreads), 0, s, a);
}

void launch_resident_root_list(hipStream_t s, const uint32_t* root_slot, uint32_t slots, uint64_t* codes,
                               uint32_t* list_slots, uint64_t* count)
{
    TimedLaunch t(kKResidentInsert, s);
    hipLaunchKernelGGL(k_resident_root_list, dim3(1), dim3(kRootThreads), 0, s, root_slot, slots, codes, list_slots,
                       count);
}

void launch_merge_insert(hipStream_t s, const MergeSource* d_sources, uint32_t num_sources, uint64_t total,
                         const uint32_t* child_slot_of, const uint32_t* child_rank_of_slot, uint8_t* masks_out,
                         uint32_t* children_out, uint64_t* table_keys, uint32_t* table_rep, const uint32_t* table_log2,
                         uint32_t* slot_of, BuildCounters* counters, int ordered, uint64_t* list_keys,
                         uint32_t* list_slots)
{
    if (total == 0) return;
    TimedLaunch t(kKMergeInsert, s);
    hipLaunchKernelGGL(k_merge_insert, dim3(grid_for((total + kThreads - 1) / kThreads, 8)), dim3(kThreads), 0, s,
                       d_sources, num_sources, total, child_slot_of, child_rank_of_slot, masks_out, children_out,
                       table_keys, table_rep, table_log2, slot_of, counters, ordered, list_keys, list_slots);
}

}  // namespace svdag
