// svdag_api.hip -- C-ABI driver of the gfx950 DAG build (include/svdag.h).
//
// Host orchestration of DAGCompression::CreateDAG / MergeDAGs / CreateFinalDAG
// (src/DAGCompression.cu:1003-1107, 1447-1548, 1550-1610).  The reference
// synchronises the host ~20 times per level (event timers, counts, frees,
// src/DAGCompression.cu:252-268, 419, 457, 894); here every count lives in
// device memory, kernels size themselves from it, and a build synchronises once.
// Scratch comes from a per-context arena (the FPool counterpart,
// src/DAGCompression.cu:32-269) sized by svdag_scratch_bytes; results live in
// blocks of a per-context pool, owned by the returned svdag_dag until freed.
#include "svdag_device.hpp"
#include "svdag_internal.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

using namespace svdag;

// Fixed device scratch of svdag_create_dag_noalloc, one copy per GPU that the
// library runs on, held until the process ends.  In host code this name is a
// host-side stand-in, not a device address: host code uses only the address
// that hipGetSymbolAddress returns for the current device (fixed_scratch).
static __device__ __align__(256) char scratch[SVDAG_NOALLOC_SCRATCH_BYTES];

namespace {

const char* const kKernelNames[kNumKernelIds] = {
    "sort_histogram", "sort_count",  "sort_scan",     "sort_scatter_keys", "sort_scatter_pairs", "fill",
    "scan_counts",    "leaf_reduce", "leaf_compact",      "table_clear",        "leaf_insert",
    "rank_of_slot",   "level_count", "level_build",   "level_check",       "extract_count",      "level_extract",
    "final_count",    "final_positions", "final_write", "merge_insert",      "sketch",             "level_tail",
    "slot_list",      "query_voxels",    "vox_setup",   "vox_count",         "vox_write",
    "trace_paths",    "resident_insert", "sort_scatter_top", "seg_count",   "seg_scatter",
    "pair_histogram", "pair_count"};

// Event pairs around the launches of the kernels selected by `mask`; resolved
// into per-kernel totals once the stream has drained.
struct EventRecorder final : KernelRecorder {
    uint64_t mask = 0;
    std::vector<hipEvent_t> pool;
    struct Pending {
        int id;
        hipEvent_t begin, end;
    };
    std::vector<Pending> pending;
    hipEvent_t open[kNumKernelIds] = {};
    double total_ms[kNumKernelIds] = {};
    uint64_t launches[kNumKernelIds] = {};

    hipEvent_t take()
    {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        return e;
    }
    void begin(hipStream_t s, int id) override
    {
        if (!((mask >> id) & 1) || open[id]) return;
        hipEvent_t e = take();
        if (e && hipEventRecord(e, s) == hipSuccess) open[id] = e;
        else if (e) pool.push_back(e);
    }
    void end(hipStream_t s, int id) override
    {
        if (!open[id]) return;
        hipEvent_t e = take();
        if (e && hipEventRecord(e, s) == hipSuccess) pending.push_back({id, open[id], e});
        else {
            pool.push_back(open[id]);
            if (e) pool.push_back(e);
        }
        open[id] = nullptr;
    }
    // The stream has drained: every pending event is complete.
    void resolve(bool keep)
    {
        for (const Pending& p : pending) {
            float ms = 0.f;
            if (keep && hipEventElapsedTime(&ms, p.begin, p.end) == hipSuccess) {
                total_ms[p.id] += ms;
                ++launches[p.id];
            }
            pool.push_back(p.begin);
            pool.push_back(p.end);
        }
        pending.clear();
        (void)hipGetLastError();
    }
    void reset()
    {
        for (int i = 0; i < kNumKernelIds; ++i) {
            total_ms[i] = 0;
            launches[i] = 0;
        }
    }
    ~EventRecorder() override
    {
        resolve(false);
        for (hipEvent_t e : pool) (void)hipEventDestroy(e);
        for (hipEvent_t e : open)
            if (e) (void)hipEventDestroy(e);
    }
};

thread_local KernelRecorder* t_recorder = nullptr;

}  // namespace

namespace svdag {
KernelRecorder*& active_recorder() { return t_recorder; }
}  // namespace svdag

struct svdag_context {
    hipStream_t stream = nullptr;
    int device = 0;
    char* arena = nullptr;
    uint64_t arena_cap = 0;
    // Source table of the last gathered build (kept until its copy is done).
    std::vector<GatherSource> gather_src;
    // Result blocks (DAG storage, final words): reused across calls, returned
    // by svdag_free_dag / svdag_free_words, released by svdag_free_all.
    std::multimap<uint64_t, void*> free_blocks;
    std::unordered_map<void*, uint64_t> block_size;
    uint64_t pool_bytes = 0;
    BuildCounters* h_counters = nullptr;  // pinned
    int check_collisions = 0;
    uint64_t budget = 0;
    uint32_t load_shift = 1;
    int radix_bits = 9;
    int sort_ipt = kSortIptSmall;
    uint32_t seg_first_bits = 0;
    int top_first = kTopFirstAuto;
    bool table_estimate = true;
    uint64_t vox_stage_max = kVoxStageMaxBytes;
    int host_threads = 1;
    // Stage markers of the work in flight (bounded completion wait).
    static constexpr int kMaxMarks = 64;
    hipEvent_t marks[kMaxMarks] = {};
    const char* mark_names[kMaxMarks] = {};
    int num_marks = 0;
    double wait_seconds = 120.0;
    EventRecorder recorder;
    // Second stream (default priority) for work that may overlap the main
    // chain; it forks from and joins back to `stream` through the two events.
    hipStream_t side = nullptr;
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
};

namespace {

thread_local std::string g_error;

int set_error(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_error = buf;
    return code;
}

#define HIP_TRY(expr)                                                                                    \
    do {                                                                                                 \
        const hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess) return set_error(SVDAG_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

#define TRY(expr)                    \
    do {                             \
        const int rc_ = (expr);      \
        if (rc_ != 0) return rc_;    \
    } while (0)

// Records a named marker after the work enqueued so far.
// Builds of at least this many keys order their unique leaves on the side stream.
constexpr uint64_t kOverlapMinKeys = 1ull << 24;

// The side stream and its events, created on first use; false if the runtime
// refuses them (the work then stays on the main stream).
bool ensure_side(svdag_context* ctx)
{
    if (ctx->side) return true;
    hipStream_t s = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) == hipSuccess &&
        hipEventCreateWithFlags(&a, hipEventDisableTiming) == hipSuccess &&
        hipEventCreateWithFlags(&b, hipEventDisableTiming) == hipSuccess) {
        ctx->side = s;
        ctx->fork_ev = a;
        ctx->join_ev = b;
        return true;
    }
    if (b) (void)hipEventDestroy(b);
    if (a) (void)hipEventDestroy(a);
    if (s) (void)hipStreamDestroy(s);
    (void)hipGetLastError();
    return false;
}

void mark(svdag_context* ctx, const char* name)
{
    if (ctx->num_marks >= svdag_context::kMaxMarks || !ctx->marks[ctx->num_marks]) return;
    if (hipEventRecord(ctx->marks[ctx->num_marks], ctx->stream) == hipSuccess) ctx->mark_names[ctx->num_marks++] = name;
}

// Waits for the stream with a deadline instead of blocking without bound; on
// expiry the error names the first stage whose marker has not completed.
int finish(svdag_context* ctx, const char* what)
{
    const auto start = std::chrono::steady_clock::now();
    for (;;) {
        const hipError_t e = hipStreamQuery(ctx->stream);
        if (e == hipSuccess) break;
        if (e != hipErrorNotReady) {
            ctx->num_marks = 0;
            ctx->recorder.resolve(false);
            return set_error(SVDAG_ERR_HIP, "%s failed: %s", what, hipGetErrorString(e));
        }
        const double waited = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
        if (waited > ctx->wait_seconds) {
            const char* stalled = "after the last marker";
            for (int i = 0; i < ctx->num_marks; ++i)
                if (hipEventQuery(ctx->marks[i]) != hipSuccess) {
                    stalled = ctx->mark_names[i];
                    break;
                }
            (void)hipGetLastError();
            ctx->num_marks = 0;
            return set_error(SVDAG_ERR_INTERNAL, "%s: device work did not complete within %.0f s (stalled in: %s)", what,
                             ctx->wait_seconds, stalled);
        }
        // Short calls are noticed promptly (a sleep rounds up to tens of
        // microseconds); long ones poll every 200 us.
        if (waited < 0.05) std::this_thread::yield();
        else std::this_thread::sleep_for(std::chrono::microseconds(200));
    }
    ctx->num_marks = 0;
    ctx->recorder.resolve(true);
    return 0;
}

constexpr uint64_t kAlign = 256;
inline uint64_t align_up(uint64_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

struct Carver {
    uint64_t off = 0;
    uint64_t take(uint64_t bytes)
    {
        const uint64_t o = off;
        off += align_up(std::max<uint64_t>(bytes, 1));
        return o;
    }
};

inline uint64_t pow8_capped(int e, uint64_t cap)
{
    uint64_t v = 1;
    for (int i = 0; i < e; ++i) {
        if (v > cap / 8) return cap;
        v *= 8;
    }
    return std::min(v, cap);
}

inline uint32_t table_lg(uint64_t count, uint32_t shift)
{
    return std::max(8u, ceil_log2_u64(std::max<uint64_t>(count, 1) << shift));
}

// Scratch layout of one build over n keys at `depth`.  Only [zero_begin,
// zero_end) (counters and sort histograms) must be zero at the start; every
// other buffer is fully written before it is read.
struct BuildLayout {
    uint64_t total = 0;
    uint64_t zero_begin = 0, zero_end = 0;
    uint64_t counters = 0, key_sort_zero = 0, leaf_sort_zero = 0, sketch = 0;
    uint64_t sort_counts = 0, tile_counts = 0, tile_offsets = 0;
    uint64_t keys_a = 0, keys_b = 0;
    uint64_t lp_code = 0, lp_mask = 0, lp_slot = 0, leaf_table = 0, leaf_rank = 0;
    uint64_t reg_code = 0, reg_mask = 0;  // leaf-reduce regions, over the later stages' buffers
    uint64_t list_keys[2] = {}, list_slots[2] = {}, sorted_slots = 0;
    // Levels: build outputs (codes, slots, hashes) ping-pong between
    // consecutive level builds; every hashed level keeps its own records
    // (masks, children) and table until its extract, so that all level builds
    // can run before the first extract (which waits for the leaf order).
    uint64_t codes[2] = {}, slots[2] = {}, hashes[2] = {};
    uint64_t masks[SVDAG_MAX_LEVELS] = {}, children[SVDAG_MAX_LEVELS] = {};
    uint64_t tkeys[SVDAG_MAX_LEVELS] = {}, trep[SVDAG_MAX_LEVELS] = {}, trank[2] = {};
    uint64_t level_nodes[SVDAG_MAX_LEVELS] = {}, level_slots_of[SVDAG_MAX_LEVELS] = {};
    uint64_t tail_codes[2] = {}, tail_ranks[2] = {}, tail_masks = 0, tail_children = 0, tail_hash = 0;
    uint64_t n = 0, leaf_bound = 0, level_bound = 0, leaf_slots = 0, level_slots = 0;
    uint32_t group_bits = 0;
    SortPlan key_plan{}, leaf_plan{};
    // Plain builds of 64-bit keys may sort top digit first
    // (launch_top_digit_sort): the 32-bit bucket-local keys over keys_b, the
    // bucket starts and the segmented sort's tables over the buffers from
    // lp_code on (nothing there is live before the leaf reduce).
    bool top_first = false;
    uint32_t top_shift = 0, top_bits = 0;
    uint64_t bucket_start = 0, seg_tiles = 0, seg_ct = 0, seg_counts = 0, seg_totals = 0, seg_starts = 0,
             seg_ntiles = 0;
};

BuildLayout plan_build(uint64_t n, int depth, int key_bytes, uint32_t load_shift, bool batched, int radix_bits,
                       int sort_ipt, int top_first = kTopFirstNever)
{
    BuildLayout L;
    L.n = n;
    L.leaf_bound = std::min<uint64_t>(n, pow8_capped(depth - 2, n));
    L.level_bound = std::min<uint64_t>(L.leaf_bound, pow8_capped(depth - 3, n));
    L.leaf_slots = 1ull << table_lg(L.leaf_bound, load_shift);
    L.level_slots = (1ull << table_lg(L.level_bound, load_shift)) + kTableOverflow;
    // Keys only need grouping by code >> (6 + group_bits): the leaf reduce ORs
    // the low 6 bits into a mask and indexes its table by the next group_bits,
    // so the sort skips both.  It keeps at least one bit so that its first read
    // (range check, chunk prefix) always runs; at depth <= 5 the whole key list
    // is then one group.
    const uint32_t code_bits = uint32_t(3 * depth - 6);
    L.group_bits = std::min(kLeafGroupBits, code_bits);
    L.key_plan = make_sort_plan(int(6 + std::min(L.group_bits, code_bits - 1)), 3 * depth, radix_bits, sort_ipt);
    L.leaf_plan = make_sort_plan(0, 64);
    const uint64_t stage_tiles = std::max({tiles(n, kPackTileItems), tiles(L.leaf_bound, kBuildTileItems),
                                           tiles(L.level_slots, kExtractTileItems),
                                           tiles(L.leaf_slots, kListTileItems), uint64_t(1)});
    Carver c;
    L.zero_begin = c.off;
    L.counters = c.take(sizeof(BuildCounters));
    L.key_sort_zero = c.take(4 * sort_zero_words(L.key_plan));
    L.leaf_sort_zero = c.take(4 * sort_zero_words(L.leaf_plan));
    L.sketch = c.take(4 * kSketchRegs);
    L.zero_end = c.off;
    L.sort_counts = c.take(4 * sort_count_words(n));
    L.tile_counts = c.take(4 * stage_tiles);
    L.tile_offsets = c.take(8 * stage_tiles);
    const int kb = batched ? 8 : key_bytes;
    L.keys_a = c.take(uint64_t(kb) * n);
    L.keys_b = c.take(uint64_t(kb) * n);
    L.lp_code = c.take(8 * L.leaf_bound);
    L.lp_mask = c.take(8 * L.leaf_bound);
    L.lp_slot = c.take(4 * L.leaf_bound);
    L.leaf_table = c.take(8 * L.leaf_slots);
    L.leaf_rank = c.take(4 * L.leaf_slots);
    for (int i = 0; i < 2; ++i) {
        L.list_keys[i] = c.take(8 * L.leaf_bound);
        L.list_slots[i] = c.take(4 * L.leaf_bound);
    }
    L.sorted_slots = c.take(4 * L.leaf_bound);
    for (int i = 0; i < 2; ++i) {
        L.codes[i] = c.take(8 * L.level_bound);
        L.slots[i] = c.take(4 * L.level_bound);
        L.hashes[i] = c.take(8 * L.level_bound);
    }
    // Hashed levels tail_top + 1 .. depth - 3: at most min(level bound, 8^l)
    // nodes each.
    const int tail_top = std::min(depth - 3, kTailTopLevel);
    uint64_t widest_slots = 1;
    for (int l = tail_top + 1; l <= depth - 3; ++l) {
        L.level_nodes[l] = std::min<uint64_t>(L.level_bound, pow8_capped(l, L.level_bound));
        L.level_slots_of[l] = (1ull << table_lg(L.level_nodes[l], load_shift)) + kTableOverflow;
        widest_slots = std::max(widest_slots, L.level_slots_of[l]);
        L.masks[l] = c.take(L.level_nodes[l]);
        L.children[l] = c.take(32 * L.level_nodes[l]);
        L.tkeys[l] = c.take(8 * L.level_slots_of[l]);
        L.trep[l] = c.take(4 * L.level_slots_of[l]);
    }
    for (int i = 0; i < 2; ++i) L.trank[i] = c.take(4 * widest_slots);
    for (int i = 0; i < 2; ++i) {
        L.tail_codes[i] = c.take(8 * kTailMaxParents);
        L.tail_ranks[i] = c.take(4 * kTailMaxParents);
    }
    L.tail_masks = c.take(kTailMaxParents);
    L.tail_children = c.take(32 * kTailMaxParents);
    L.tail_hash = c.take(8 * 8 * kTailMaxParents);
    L.total = c.off;
    // The leaf-reduce regions are read only by the compaction, before any of
    // the buffers from the leaf table on is written.
    const uint64_t region_bytes = 8 * tiles(n, kPackTileItems) * kLeafRegionItems;
    L.reg_code = L.leaf_table;
    L.reg_mask = L.reg_code + region_bytes;
    L.total = std::max(L.total, L.reg_mask + region_bytes);
    // Top digit first: a digit of radix_bits at the top of the code, whose
    // bucket-local keys fit 32 bits, with at most kSegMaxBits sorted below it.
    const uint32_t lo = L.key_plan.shift[0];
    const uint32_t top_shift = uint32_t(3 * depth - radix_bits);
    L.top_first = top_first != kTopFirstNever && !batched && key_bytes == 8 && L.key_plan.passes >= 2 &&
                  (top_first == kTopFirstAlways || n >= kTopFirstMinKeys) && top_shift <= 32 && top_shift > lo &&
                  top_shift - lo <= kSegMaxBits;
    if (L.top_first) {
        L.top_shift = top_shift;
        L.top_bits = uint32_t(radix_bits);
        Carver t;
        t.off = L.lp_code;
        const uint64_t tb = seg_tile_bound(n, kRadix);
        L.bucket_start = t.take(8 * (kRadix + 1));
        L.seg_tiles = t.take(sizeof(SegTile) * tb);
        L.seg_ct = t.take(4 * (kRadix + 1));
        L.seg_counts = t.take(4 * (tb << kMaxSegDigitBits));
        L.seg_totals = t.take(4 * (uint64_t(kRadix) << kMaxSegDigitBits));
        L.seg_starts = t.take(4 * (uint64_t(kRadix) << kMaxSegDigitBits));
        L.seg_ntiles = t.take(8);
        L.total = std::max(L.total, t.off);
    }
    return L;
}

// Device result storage: per level hashes, children, masks at capacity
// min(8^l, leaf bound); leaves at the leaf bound.
struct DagStorage {
    uint64_t total = 0;
    uint64_t hashes[SVDAG_MAX_LEVELS] = {}, children[SVDAG_MAX_LEVELS] = {}, masks[SVDAG_MAX_LEVELS] = {};
    uint64_t leaves = 0;
};

DagStorage plan_storage(int num_levels, const uint64_t* caps, uint64_t leaf_cap)
{
    DagStorage S;
    Carver c;
    for (int l = 0; l < num_levels; ++l) {
        S.hashes[l] = c.take(8 * caps[l]);
        S.children[l] = c.take(32 * caps[l]);
        S.masks[l] = c.take(caps[l]);
    }
    S.leaves = c.take(8 * leaf_cap);
    S.total = c.off;
    return S;
}

void pool_trim(svdag_context* ctx);

// Device memory runs out, or the context's budget would be passed: like the
// reference's allocator (src/Memory.cpp:56-70, with the pool's fallback at
// src/DAGCompression.cu:85-89, 178-189), the cached free result blocks are
// released and the allocation is tried once more; if that fails too, the call
// returns SVDAG_ERR_MEMORY (the reference aborts).
int ensure_arena(svdag_context* ctx, uint64_t bytes)
{
    if (ctx->budget && bytes > ctx->budget)
        return set_error(SVDAG_ERR_MEMORY, "build needs %llu scratch bytes, budget is %llu",
                         (unsigned long long)bytes, (unsigned long long)ctx->budget);
    if (bytes <= ctx->arena_cap) return 0;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (ctx->arena) HIP_TRY(hipFree(ctx->arena));
    ctx->arena = nullptr;
    ctx->arena_cap = 0;
    // Some headroom for the next, slightly larger call, within the budget.
    uint64_t cap = align_up(bytes + bytes / 8);
    if (ctx->budget) cap = std::max(align_up(bytes), std::min(cap, ctx->budget / kAlign * kAlign));
    if (ctx->budget && cap + ctx->pool_bytes > ctx->budget) pool_trim(ctx);
    if (ctx->budget && cap + ctx->pool_bytes > ctx->budget)
        return set_error(SVDAG_ERR_MEMORY, "%llu bytes of scratch and %llu of results exceed the budget of %llu",
                         (unsigned long long)cap, (unsigned long long)ctx->pool_bytes, (unsigned long long)ctx->budget);
    for (int attempt = 0; attempt < 2; ++attempt) {
        if (hipMalloc(reinterpret_cast<void**>(&ctx->arena), cap) == hipSuccess) {
            ctx->arena_cap = cap;
            return 0;
        }
        (void)hipGetLastError();
        ctx->arena = nullptr;
        if (attempt == 0) pool_trim(ctx);
    }
    return set_error(SVDAG_ERR_MEMORY, "cannot allocate %llu bytes of device scratch", (unsigned long long)cap);
}

// Result allocation from the context's block pool: a free block of at least
// the request (and at most twice it) is reused, otherwise a new one is made.
int device_alloc(svdag_context* ctx, void** p, uint64_t bytes)
{
    *p = nullptr;
    bytes = bytes <= (1ull << 20) ? align_up(std::max<uint64_t>(bytes, 1)) : (bytes + (1ull << 20) - 1) >> 20 << 20;
    auto it = ctx->free_blocks.lower_bound(bytes);
    if (it != ctx->free_blocks.end() && it->first <= 2 * bytes) {
        *p = it->second;
        ctx->free_blocks.erase(it);
        return 0;
    }
    if (ctx->budget && ctx->arena_cap + ctx->pool_bytes + bytes > ctx->budget) pool_trim(ctx);
    if (ctx->budget && ctx->arena_cap + ctx->pool_bytes + bytes > ctx->budget)
        return set_error(SVDAG_ERR_MEMORY, "result block of %llu bytes exceeds the memory budget", (unsigned long long)bytes);
    for (int attempt = 0; attempt < 2; ++attempt) {
        if (hipMalloc(p, bytes) == hipSuccess) {
            ctx->block_size[*p] = bytes;
            ctx->pool_bytes += bytes;
            return 0;
        }
        (void)hipGetLastError();
        *p = nullptr;
        if (attempt == 0) pool_trim(ctx);
    }
    return set_error(SVDAG_ERR_MEMORY, "cannot allocate %llu bytes of device memory", (unsigned long long)bytes);
}

// Returns a result block to the pool (stream order keeps later reuse safe).
// A block the pool does not know is freed directly.
void device_release(svdag_context* ctx, void* p)
{
    if (!p) return;
    const auto it = ctx->block_size.find(p);
    if (it == ctx->block_size.end()) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipFree(p);
        return;
    }
    ctx->free_blocks.emplace(it->second, p);
}

// Frees the cached free result blocks (work in flight may still use them
// until the context's streams drain).
void pool_trim(svdag_context* ctx)
{
    if (ctx->free_blocks.empty()) return;
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->side) (void)hipStreamSynchronize(ctx->side);
    for (auto& kv : ctx->free_blocks) {
        (void)hipFree(kv.second);
        ctx->pool_bytes -= ctx->block_size[kv.second];
        ctx->block_size.erase(kv.second);
    }
    ctx->free_blocks.clear();
}

// Host storage of DAGs copied to the host: pinned blocks (copies run at the
// link's speed, and their pages are resident), kept in a process-wide pool when
// freed and reused for later copies; released by svdag_free_all.
std::mutex g_host_mutex;
std::multimap<uint64_t, void*> g_host_free;
std::unordered_map<void*, uint64_t> g_host_size;

void* host_alloc(uint64_t bytes)
{
    bytes = (std::max<uint64_t>(bytes, 1) + 4095) / 4096 * 4096;
    {
        std::lock_guard<std::mutex> lock(g_host_mutex);
        auto it = g_host_free.lower_bound(bytes);
        if (it != g_host_free.end() && it->first <= 2 * bytes) {
            void* p = it->second;
            g_host_free.erase(it);
            return p;
        }
    }
    void* p = nullptr;
    if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_host_mutex);
    g_host_size[p] = bytes;
    return p;
}

void host_release(void* p)
{
    if (!p) return;
    std::lock_guard<std::mutex> lock(g_host_mutex);
    const auto it = g_host_size.find(p);
    if (it == g_host_size.end()) {
        std::free(p);  // not a pool block
        return;
    }
    g_host_free.emplace(it->second, p);
}

void host_trim()
{
    std::lock_guard<std::mutex> lock(g_host_mutex);
    for (auto& kv : g_host_free) {
        (void)hipHostFree(kv.second);
        g_host_size.erase(kv.second);
    }
    g_host_free.clear();
}

void empty_dag(svdag_dag* out, int num_levels, int location)
{
    std::memset(out, 0, sizeof(*out));
    out->num_levels = num_levels;
    out->location = location;
}

// A batched build: num_chunks chunks back to back.  Their offsets come from
// the host (h_offsets, uploaded by the build) or already sit in device memory
// (d_offsets, a captured build's: checked and counted on the device).
struct BatchInfo {
    const uint64_t* h_offsets;
    uint64_t num_chunks;
    int chunk_depth;
    const uint64_t* d_offsets = nullptr;
};

// Returned by build() when a table sized from an estimate overflowed; the
// caller repeats the build with bound-sized tables.
constexpr int kRetryWithBounds = -100;

// Memory of a build that allocates nothing: scratch in place of the context's
// arena, the caller's buffer in place of a pooled result block.  Such a build
// also keeps to the context's stream (no side stream, which is created on
// first use) and times no kernels (timing events are created on demand).
struct BuildResult;
struct FixedMemory {
    char* arena;
    uint64_t arena_bytes;
    char* storage;
    uint64_t storage_bytes;
    // Non-null: the build is being captured into a graph.  It enqueues its
    // work and the counter read-back only, and leaves here what completing a
    // replay needs (complete_build).
    BuildResult* captured = nullptr;
};

// Where a build's result lives, for turning its counters into a svdag_dag.
struct BuildResult {
    int levels = 0;
    char* O = nullptr;
    DagStorage S;
    uint64_t* leaves = nullptr;
    uint64_t code_mask = 0;
    bool batched = false;
};

// The result of a build whose work has finished, from its counters: fills
// `out` (arrays in the result block) or returns the error the counters report
// (kRetryWithBounds when a table sized from an estimate overflowed).
int complete_build(const BuildResult& R, const BuildCounters& h, bool estimate, svdag_dag* out)
{
    if (!R.batched && (h.key_or & ~R.code_mask))
        return set_error(SVDAG_ERR_ARGUMENT, "a key is >= 2^(3*depth) for depth %d", R.levels + 2);
    if (h.error & kErrKeyRange) return set_error(SVDAG_ERR_ARGUMENT, "a chunk-local key is >= 2^(3*chunk_depth)");
    if (h.error & kErrOffsets)
        return set_error(SVDAG_ERR_ARGUMENT, "chunk offsets do not start at 0, decrease, or exceed the key capacity");
    if ((h.error & kErrTableOverflow) && estimate) return kRetryWithBounds;
    if (h.error & kErrTableOverflow) return set_error(SVDAG_ERR_INTERNAL, "dedup table probe overflow");
    if (h.error & kErrInternal) return set_error(SVDAG_ERR_INTERNAL, "leaf reduce region overflow");
    empty_dag(out, R.levels, SVDAG_DEVICE);
    out->num_keys = h.n;
    if (h.n == 0) return 0;
    if (h.uniq[0] != 1) return set_error(SVDAG_ERR_INTERNAL, "root level holds %llu nodes", (unsigned long long)h.uniq[0]);
    out->leaves = R.leaves;
    out->num_leaves = h.unique_leaves;
    out->leaf_positions = h.leaf_positions;
    out->hash_collisions = h.collisions;
    for (int l = 0; l < R.levels; ++l) {
        out->levels[l].hashes = reinterpret_cast<uint64_t*>(R.O + R.S.hashes[l]);
        out->levels[l].children = reinterpret_cast<uint32_t*>(R.O + R.S.children[l]);
        out->levels[l].masks = reinterpret_cast<uint8_t*>(R.O + R.S.masks[l]);
        out->levels[l].num_nodes = h.uniq[l];
        out->level_words[l] = h.words[l];
        out->pre_dedup[l] = h.pre[l];
    }
    return 0;
}

// Scratch of a build beyond plan_build's layout: a batched build's device
// offsets and, when it sorts inside each chunk, the segmented sort's tables.
struct BatchLayout {
    uint64_t total = 0;  // bytes from BuildLayout::total on
    uint64_t offsets = 0, seg_tiles = 0, seg_ct = 0, seg_counts = 0, seg_totals = 0, seg_starts = 0, seg_ntiles = 0;
    bool segmented = false;
    uint32_t seg_lo = 0, seg_chunk_bits = 0;
};

BatchLayout plan_batch(const BuildLayout& L, uint64_t n, const BatchInfo* batch)
{
    BatchLayout B;
    if (!batch) return B;
    Carver c;
    c.off = L.total;
    B.offsets = c.take(8 * (batch->num_chunks + 1));
    // Batched builds whose chunk-local sort spans at most kSegMaxBits bits sort
    // inside each chunk (the chunk order is the input order); the tiles are
    // planned on the device from the offsets.
    B.seg_lo = uint32_t(L.key_plan.shift[0]);
    B.seg_chunk_bits = uint32_t(3 * batch->chunk_depth);
    B.segmented = L.key_plan.passes > 0 && B.seg_chunk_bits > B.seg_lo && B.seg_chunk_bits - B.seg_lo <= kSegMaxBits;
    if (B.segmented) {
        const uint64_t tb = seg_tile_bound(n, batch->num_chunks);
        B.seg_tiles = c.take(sizeof(SegTile) * tb);
        B.seg_ct = c.take(4 * (batch->num_chunks + 1));
        B.seg_counts = c.take(4 * (tb << kMaxSegDigitBits));
        B.seg_totals = c.take(4 * (batch->num_chunks << kMaxSegDigitBits));
        B.seg_starts = c.take(4 * (batch->num_chunks << kMaxSegDigitBits));
        B.seg_ntiles = c.take(8);
    }
    B.total = c.off - L.total;
    return B;
}

// Result storage of a build over n keys at `depth` (the layout of plan_storage).
uint64_t result_bytes(uint64_t n, int depth)
{
    const int levels = depth - 2;
    const uint64_t leaf_bound = std::min<uint64_t>(n, pow8_capped(depth - 2, n));
    uint64_t caps[SVDAG_MAX_LEVELS];
    for (int l = 0; l < levels; ++l) caps[l] = pow8_capped(l, leaf_bound);
    return plan_storage(levels, caps, leaf_bound).total;
}

template <typename K>
int build(svdag_context* ctx, const K* d_keys, uint64_t n, int depth, svdag_dag* out, const BatchInfo* batch,
          bool estimate, const FixedMemory* fixed = nullptr)
{
    hipStream_t st = ctx->stream;
    const int levels = depth - 2;
    empty_dag(out, levels, SVDAG_DEVICE);
    out->num_keys = n;
    if (n == 0) return 0;  // DAGCompression.cu:1010-1013

    const BuildLayout L = plan_build(n, depth, int(sizeof(K)), ctx->load_shift, batch != nullptr, ctx->radix_bits,
                                     ctx->sort_ipt, ctx->top_first);
    const BatchLayout B = plan_batch(L, n, batch);
    const uint64_t scratch_need = L.total + B.total;
    if (fixed && scratch_need > fixed->arena_bytes)
        return set_error(SVDAG_ERR_MEMORY, "build needs %llu scratch bytes, the fixed scratch holds %llu",
                         (unsigned long long)scratch_need, (unsigned long long)fixed->arena_bytes);
    if (!fixed) TRY(ensure_arena(ctx, scratch_need));
    char* A = fixed ? fixed->arena : ctx->arena;
    BuildCounters* c = reinterpret_cast<BuildCounters*>(A + L.counters);
    unsigned long long* err = reinterpret_cast<unsigned long long*>(&c->error);
    uint32_t* sort_counts = reinterpret_cast<uint32_t*>(A + L.sort_counts);
    uint32_t* tile_counts = reinterpret_cast<uint32_t*>(A + L.tile_counts);
    uint64_t* tile_offsets = reinterpret_cast<uint64_t*>(A + L.tile_offsets);

    uint64_t caps[SVDAG_MAX_LEVELS];
    for (int l = 0; l < levels; ++l) caps[l] = pow8_capped(l, L.leaf_bound);
    const DagStorage S = plan_storage(levels, caps, L.leaf_bound);
    void* storage = nullptr;
    if (fixed) {
        if (S.total > fixed->storage_bytes)
            return set_error(SVDAG_ERR_MEMORY, "result needs %llu bytes, the caller's storage holds %llu",
                             (unsigned long long)S.total, (unsigned long long)fixed->storage_bytes);
        storage = fixed->storage;
    } else {
        TRY(device_alloc(ctx, &storage, S.total));
    }
    char* O = static_cast<char*>(storage);
    const bool capture = fixed && fixed->captured;
    bool forked = false;  // work is in flight on the side stream
    auto fail = [&](int code) {
        if (forked) (void)hipStreamSynchronize(ctx->side);
        if (!fixed) device_release(ctx, storage);
        out->storage = nullptr;
        return code;
    };

    // The zeroed region holds the counters.  The key count: given (set by the
    // same fill), or (device offsets) checked and read on the device.
    const uint64_t n_word = (L.counters + offsetof(BuildCounters, n) - L.zero_begin) / 8;
    const uint64_t* d_off = nullptr;
    if (batch && batch->d_offsets) {
        launch_fill_u64(st, reinterpret_cast<uint64_t*>(A + L.zero_begin), (L.zero_end - L.zero_begin) / 8, 0);
        d_off = batch->d_offsets;
        launch_batch_count(st, d_off, batch->num_chunks, n, &c->n, err);
    } else {
        launch_fill_u64(st, reinterpret_cast<uint64_t*>(A + L.zero_begin), (L.zero_end - L.zero_begin) / 8, 0, n_word, n);
    }

    // Sort keys on bits [0, 3*depth) (SortKeys, DAGCompression.cu:1024-1026).
    int where;
    const void* sorted;
    // Segmented sorts leave 32-bit keys local to these buckets.
    ChunkPrefix buckets;
    if (batch) {
        if (!d_off) {
            uint64_t* up = reinterpret_cast<uint64_t*>(A + B.offsets);
            if (hipMemcpyAsync(up, batch->h_offsets, 8 * (batch->num_chunks + 1), hipMemcpyHostToDevice, st) != hipSuccess)
                return fail(set_error(SVDAG_ERR_HIP, "offset upload failed"));
            d_off = up;
        }
        // The sort's first reads put each chunk's index above its chunk-local keys.
        ChunkPrefix prefix;
        prefix.offsets = d_off;
        prefix.num_chunks = batch->num_chunks;
        prefix.chunk_bits = uint32_t(3 * batch->chunk_depth);
        prefix.err = err;
        uint64_t* ka = reinterpret_cast<uint64_t*>(A + L.keys_a);
        uint64_t* kb = reinterpret_cast<uint64_t*>(A + L.keys_b);
        if (B.segmented) {
            buckets.offsets = d_off;
            buckets.num_chunks = batch->num_chunks;
            buckets.chunk_bits = B.seg_chunk_bits;
            SegScratch sc;
            sc.tiles = reinterpret_cast<SegTile*>(A + B.seg_tiles);
            sc.chunk_tiles = reinterpret_cast<uint32_t*>(A + B.seg_ct);
            sc.counts = reinterpret_cast<uint32_t*>(A + B.seg_counts);
            sc.totals = reinterpret_cast<uint32_t*>(A + B.seg_totals);
            sc.starts = reinterpret_cast<uint32_t*>(A + B.seg_starts);
            sc.ntiles = reinterpret_cast<uint64_t*>(A + B.seg_ntiles);
            where = launch_segmented_sort(st, reinterpret_cast<const uint64_t*>(d_keys),
                                          reinterpret_cast<uint32_t*>(ka), reinterpret_cast<uint32_t*>(kb), d_off,
                                          batch->num_chunks, &c->n, n, sc, B.seg_lo, B.seg_chunk_bits, err,
                                          ctx->seg_first_bits);
        } else {
            where = launch_sort<uint64_t>(st, L.key_plan, reinterpret_cast<const uint64_t*>(d_keys), nullptr, ka, kb,
                                          nullptr, nullptr, nullptr, nullptr, &c->n, n,
                                          reinterpret_cast<uint32_t*>(A + L.key_sort_zero), sort_counts, &c->key_or,
                                          prefix);
        }
        sorted = where == 0 ? static_cast<const void*>(ka) : static_cast<const void*>(kb);
    } else {
        K* ka = reinterpret_cast<K*>(A + L.keys_a);
        K* kb = reinterpret_cast<K*>(A + L.keys_b);
        where = -1;
        if constexpr (sizeof(K) == 8) {
            if (L.top_first) {
                SegScratch sc;
                sc.tiles = reinterpret_cast<SegTile*>(A + L.seg_tiles);
                sc.chunk_tiles = reinterpret_cast<uint32_t*>(A + L.seg_ct);
                sc.counts = reinterpret_cast<uint32_t*>(A + L.seg_counts);
                sc.totals = reinterpret_cast<uint32_t*>(A + L.seg_totals);
                sc.starts = reinterpret_cast<uint32_t*>(A + L.seg_starts);
                sc.ntiles = reinterpret_cast<uint64_t*>(A + L.seg_ntiles);
                where = launch_top_digit_sort(st, d_keys, reinterpret_cast<uint32_t*>(kb),
                                              reinterpret_cast<uint32_t*>(ka), reinterpret_cast<uint32_t*>(kb),
                                              L.key_plan.ipt, L.top_shift, L.top_bits, L.key_plan.shift[0], &c->n, n,
                                              reinterpret_cast<uint32_t*>(A + L.key_sort_zero), sort_counts,
                                              &c->key_or, reinterpret_cast<uint64_t*>(A + L.bucket_start), sc, err);
                buckets.offsets = reinterpret_cast<const uint64_t*>(A + L.bucket_start);
                buckets.num_chunks = 1ull << L.top_bits;
                buckets.chunk_bits = L.top_shift;
            }
        }
        if (!L.top_first)
            where = launch_sort<K>(st, L.key_plan, d_keys, nullptr, ka, kb, nullptr, nullptr, nullptr, nullptr, &c->n,
                                   n, reinterpret_cast<uint32_t*>(A + L.key_sort_zero), sort_counts, &c->key_or);
        sorted = where == 0 ? static_cast<const void*>(ka) : static_cast<const void*>(kb);
    }
    if (where < 0) return fail(set_error(SVDAG_ERR_INTERNAL, "sort plan failed"));
    if (!capture) mark(ctx, "key sort");

    uint64_t* lp_code = reinterpret_cast<uint64_t*>(A + L.lp_code);
    uint64_t* lp_mask = reinterpret_cast<uint64_t*>(A + L.lp_mask);
    uint32_t* lp_slot = reinterpret_cast<uint32_t*>(A + L.lp_slot);
    const uint64_t code_mask = depth >= 22 ? ~0ull : ((1ull << (3 * depth)) - 1);
    // Leaf ReduceByKey (DAGCompression.cu:1044-1068).
    uint64_t* reg_code = reinterpret_cast<uint64_t*>(A + L.reg_code);
    uint64_t* reg_mask = reinterpret_cast<uint64_t*>(A + L.reg_mask);
    if (buckets.offsets || (!batch && sizeof(K) == 4))
        launch_leaf_reduce<uint32_t>(st, static_cast<const uint32_t*>(sorted), c, L.group_bits, reg_code, reg_mask,
                                     lp_code, lp_mask, L.leaf_bound, tile_counts, tile_offsets, n, depth - 3,
                                     code_mask, buckets);
    else
        launch_leaf_reduce<uint64_t>(st, static_cast<const uint64_t*>(sorted), c, L.group_bits, reg_code, reg_mask,
                                     lp_code, lp_mask, L.leaf_bound, tile_counts, tile_offsets, n, depth - 3,
                                     code_mask);

    // SortLevel(Leaves) (DAGCompression.cu:1070-1071): dedup table, then order.
    uint64_t* leaf_table = reinterpret_cast<uint64_t*>(A + L.leaf_table);
    uint32_t* leaf_rank = reinterpret_cast<uint32_t*>(A + L.leaf_rank);
    uint32_t* leaf_lg = &c->table_log2[SVDAG_MAX_LEVELS];
    uint64_t* lk0 = reinterpret_cast<uint64_t*>(A + L.list_keys[0]);
    uint64_t* lk1 = reinterpret_cast<uint64_t*>(A + L.list_keys[1]);
    uint32_t* ls0 = reinterpret_cast<uint32_t*>(A + L.list_slots[0]);
    uint32_t* ls1 = reinterpret_cast<uint32_t*>(A + L.list_slots[1]);
    uint32_t* sorted_slots = reinterpret_cast<uint32_t*>(A + L.sorted_slots);
    // The leaf table is sized from a distinct-count estimate of the masks (or,
    // on a retry, from the leaf positions, an upper bound).
    const uint64_t* leaf_size = &c->leaf_positions;
    if (estimate) {
        uint32_t* regs = reinterpret_cast<uint32_t*>(A + L.sketch);
        launch_sketch(st, lp_mask, &c->leaf_positions, regs, L.leaf_bound);
        launch_sketch_estimate(st, regs, &c->leaf_positions, &c->leaf_estimate);
        leaf_size = &c->leaf_estimate;
    }
    launch_table_clear(st, leaf_table, leaf_size, leaf_lg, ctx->load_shift, 0, L.leaf_slots);
    launch_leaf_insert(st, lp_mask, &c->leaf_positions, leaf_table, leaf_lg, lp_slot, L.leaf_bound, err);
    launch_slot_list(st, leaf_table, leaf_lg, L.leaf_slots, tile_counts, tile_offsets, &c->table_slots, lk0, ls0,
                     &c->unique_leaves);
    // The unique leaves are ordered (sorted masks, slot -> rank) on the side
    // stream for large builds: the level above the leaves needs their ranks
    // only from its extract on, so its build overlaps this sort.  Small builds
    // keep one stream (many concurrent contexts would share hardware queues).
    uint64_t* leaves_out = reinterpret_cast<uint64_t*>(O + S.leaves);
    hipStream_t ls = st;
    if (n >= kOverlapMinKeys && !fixed && ensure_side(ctx) && hipEventRecord(ctx->fork_ev, st) == hipSuccess &&
        hipStreamWaitEvent(ctx->side, ctx->fork_ev, 0) == hipSuccess) {
        ls = ctx->side;
        forked = true;
    }
    // The sort's launches are issued after the first level builds' (below):
    // the side stream already waits for the leaf dedup (fork_ev), and issuing
    // builds first keeps the main stream busy while the host issues the sort's
    // two dozen small launches.
    bool ordered = false;
    auto order_leaves = [&]() -> int {
        if (ordered) return 0;
        ordered = true;
        if (launch_sort<uint64_t>(ls, L.leaf_plan, lk0, ls0, lk1, lk0, ls1, ls0, leaves_out, sorted_slots,
                                  &c->unique_leaves, L.leaf_bound, reinterpret_cast<uint32_t*>(A + L.leaf_sort_zero),
                                  sort_counts, nullptr) < 0)
            return fail(set_error(SVDAG_ERR_INTERNAL, "leaf sort plan failed"));
        launch_rank_of_slot(ls, sorted_slots, &c->unique_leaves, leaf_rank, L.leaf_bound);
        if (forked && hipEventRecord(ctx->join_ev, ls) != hipSuccess)
            return fail(set_error(SVDAG_ERR_HIP, "event record failed"));
        return 0;
    };
    bool joined = !forked;
    // Before the first use of the leaf order.
    auto join = [&]() -> int {
        TRY(order_leaves());
        if (joined) return 0;
        joined = true;
        if (hipStreamWaitEvent(st, ctx->join_ev, 0) != hipSuccess)
            return fail(set_error(SVDAG_ERR_HIP, "stream wait failed"));
        if (!capture) mark(ctx, "leaf order");
        return 0;
    };
    if (!capture) mark(ctx, "leaf dedup");

    // Levels, bottom-up (DAGCompression.cu:1077-1092): hashed-table levels down
    // to tail_top + 1, then levels tail_top..0 (at most 512 nodes each) in one
    // workgroup.  A level build needs only the children's hashes (in child
    // order) and table slots, not their ranks, so every level build is issued
    // first; the extracts follow in order (each turns its children's slots
    // into ranks of the level below), the first one after the leaf order.
    const int tail_top = std::min(depth - 3, kTailTopLevel);
    for (int l = depth - 3; l > tail_top; --l) {
        const int cur = (depth - 3 - l) & 1;
        LevelBuildArgs a{};
        if (l == depth - 3) {
            a.child_codes = lp_code;
            a.child_slots = lp_slot;
            a.child_hash_seq = lp_mask;  // a leaf's hash is its mask
            a.child_count = &c->leaf_positions;
        } else {
            a.child_codes = reinterpret_cast<const uint64_t*>(A + L.codes[cur ^ 1]);
            a.child_slots = reinterpret_cast<const uint32_t*>(A + L.slots[cur ^ 1]);
            a.child_hash_seq = reinterpret_cast<const uint64_t*>(A + L.hashes[cur ^ 1]);
            a.child_count = &c->pre[l + 1];
        }
        a.parent_hashes = reinterpret_cast<uint64_t*>(A + L.hashes[cur]);
        a.parent_codes = reinterpret_cast<uint64_t*>(A + L.codes[cur]);
        a.parent_slots = reinterpret_cast<uint32_t*>(A + L.slots[cur]);
        a.parent_masks = reinterpret_cast<uint8_t*>(A + L.masks[l]);
        a.parent_children = reinterpret_cast<uint32_t*>(A + L.children[l]);
        a.table_keys = reinterpret_cast<uint64_t*>(A + L.tkeys[l]);
        a.table_rep = reinterpret_cast<uint32_t*>(A + L.trep[l]);
        a.table_log2 = &c->table_log2[l];
        a.tile_counts = tile_counts;
        a.offsets = tile_offsets;
        a.next_count = l > 0 ? &c->pre[l - 1] : nullptr;
        a.counters = c;
        // The table is sized from the level's node count (the build of the
        // level below, or the leaf reduce, counted it) and cleared by the
        // level's count kernel.
        const TableClearArgs clear{a.table_keys, &c->pre[l], &c->table_log2[l], ctx->load_shift, kTableOverflow,
                                   L.level_slots_of[l]};
        launch_level_build(st, a, l == depth - 3 ? L.leaf_bound : L.level_nodes[l + 1], &clear);
        if (ctx->check_collisions)
            launch_level_check(st, a.parent_slots, a.table_rep, a.parent_masks, a.parent_children, &c->pre[l],
                               &c->collisions, L.level_nodes[l]);
        // The unique-leaf sort's two dozen small launches go out once the two
        // largest level builds are queued: the main stream then has enough
        // work not to wait while the host issues them, and the sort starts
        // early enough on the side stream to end beside the builds.
        if (l == std::max(depth - 4, tail_top + 1)) TRY(order_leaves());
    }
    // Stage markers (the completion deadline names the first stage not done)
    // go between stage groups only: each event record leaves the GPU idle for
    // a few microseconds.
    if (!capture) mark(ctx, "level builds");
    TRY(join());
    const uint32_t* below = leaf_rank;  // slot -> rank map of the level below
    for (int l = depth - 3; l > tail_top; --l) {
        const int cur = (depth - 3 - l) & 1;
        LevelExtractArgs e{};
        e.table_keys = reinterpret_cast<const uint64_t*>(A + L.tkeys[l]);
        e.table_rep = reinterpret_cast<const uint32_t*>(A + L.trep[l]);
        e.table_log2 = &c->table_log2[l];
        e.rank_of_slot = reinterpret_cast<uint32_t*>(A + L.trank[cur]);
        e.parent_masks = reinterpret_cast<const uint8_t*>(A + L.masks[l]);
        e.parent_children = reinterpret_cast<const uint32_t*>(A + L.children[l]);
        e.child_rank_of_slot = below;
        e.out_hashes = reinterpret_cast<uint64_t*>(O + S.hashes[l]);
        e.out_masks = reinterpret_cast<uint8_t*>(O + S.masks[l]);
        e.out_children = reinterpret_cast<uint32_t*>(O + S.children[l]);
        e.tile_counts = tile_counts;
        e.offsets = tile_offsets;
        e.slots_scratch = &c->table_slots;
        e.q_count = &c->uniq[l];
        e.w_count = &c->words[l];
        launch_level_extract(st, e, L.level_slots_of[l]);
        below = e.rank_of_slot;
    }
    if (!capture) mark(ctx, "level extracts");
    {
        TailArgs t{};
        t.top = tail_top;
        if (tail_top == depth - 3) {
            t.child_codes = lp_code;
            t.child_slots = lp_slot;
            t.child_hashes = leaves_out;
            t.child_count = &c->leaf_positions;
        } else {
            const int cur = (depth - 3 - (tail_top + 1)) & 1;  // outputs of the build of level tail_top + 1
            t.child_codes = reinterpret_cast<const uint64_t*>(A + L.codes[cur]);
            t.child_slots = reinterpret_cast<const uint32_t*>(A + L.slots[cur]);
            t.child_hashes = reinterpret_cast<const uint64_t*>(O + S.hashes[tail_top + 1]);
            t.child_count = &c->pre[tail_top + 1];
        }
        t.child_rank_of_slot = below;
        for (int i = 0; i < 2; ++i) {
            t.codes[i] = reinterpret_cast<uint64_t*>(A + L.tail_codes[i]);
            t.ranks[i] = reinterpret_cast<uint32_t*>(A + L.tail_ranks[i]);
        }
        t.masks = reinterpret_cast<uint8_t*>(A + L.tail_masks);
        t.children = reinterpret_cast<uint32_t*>(A + L.tail_children);
        t.child_hash = reinterpret_cast<uint64_t*>(A + L.tail_hash);
        for (int l = 0; l <= tail_top; ++l) {
            t.out_hashes[l] = reinterpret_cast<uint64_t*>(O + S.hashes[l]);
            t.out_masks[l] = reinterpret_cast<uint8_t*>(O + S.masks[l]);
            t.out_children[l] = reinterpret_cast<uint32_t*>(O + S.children[l]);
        }
        t.counters = c;
        t.check_collisions = ctx->check_collisions;
        launch_level_tail(st, t);
        if (!capture) mark(ctx, "top levels");
    }
    {
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return fail(set_error(SVDAG_ERR_HIP, "kernel launch failed: %s", hipGetErrorString(e)));
    }
    if (hipMemcpyAsync(ctx->h_counters, c, sizeof(BuildCounters), hipMemcpyDeviceToHost, st) != hipSuccess)
        return fail(set_error(SVDAG_ERR_HIP, "counter read-back failed: %s", hipGetErrorString(hipGetLastError())));
    BuildResult R;
    R.levels = levels;
    R.O = O;
    R.S = S;
    R.leaves = leaves_out;
    R.code_mask = code_mask;
    R.batched = batch != nullptr;
    if (capture) {
        *fixed->captured = R;
        return 0;
    }
    mark(ctx, "counter read-back");
    // On a stall or an asynchronous error the result block goes back to the
    // pool (stream order keeps its later reuse safe).
    if (const int rc = finish(ctx, "create_dag")) return fail(rc);
    if (const int rc = complete_build(R, *ctx->h_counters, estimate, out)) return fail(rc);
    out->num_keys = n;
    out->storage = fixed ? nullptr : storage;  // the caller's buffer is not the library's to free
    return 0;
}

int check_ctx(svdag_context* ctx)
{
    if (!ctx) return set_error(SVDAG_ERR_ARGUMENT, "null context");
    HIP_TRY(hipSetDevice(ctx->device));
    active_recorder() = ctx->recorder.mask ? &ctx->recorder : nullptr;
    return 0;
}

uint64_t dag_words(const svdag_dag* d)
{
    uint64_t w = 2 * d->num_leaves;
    for (int l = 0; l < d->num_levels; ++l) w += d->level_words[l];
    return w;
}

// MergeDAGs (src/DAGCompression.cu:1447-1548) over device-resident inputs.
// Levels below the sub-DAG roots: every input's nodes of that level enter one
// shared table (children remapped, hashes reused) and are extracted in hash
// order.  Levels above: the non-empty sub-DAG roots are the children, keyed by
// their chunk index (Morton order), and the ordinary level build runs on them.
int merge_device(svdag_context* ctx, const svdag_dag* dags, uint64_t count, int k, svdag_dag* out)
{
    hipStream_t st = ctx->stream;
    const int lc_levels = dags[0].num_levels;
    const int levels = lc_levels + k;
    empty_dag(out, levels, SVDAG_DEVICE);
    std::vector<uint64_t> live;
    for (uint64_t i = 0; i < count; ++i)
        if (dags[i].num_leaves) live.push_back(i);
    if (live.empty()) return 0;
    const uint64_t nlive = live.size();

    uint64_t total_leaves = 0;
    std::vector<uint64_t> totals(lc_levels, 0);
    for (uint64_t i : live) {
        total_leaves += dags[i].num_leaves;
        for (int l = 0; l < lc_levels; ++l) totals[l] += dags[i].levels[l].num_nodes;
    }
    uint64_t max_total = 0;
    for (int l = 0; l < lc_levels; ++l) max_total = std::max(max_total, totals[l]);
    max_total = std::max<uint64_t>(max_total, 1);
    if (total_leaves >= (1ull << 30) || max_total >= (1ull << 30))
        return set_error(SVDAG_ERR_CAPACITY, "merge input exceeds 2^30 nodes per level");

    const uint32_t shift = ctx->load_shift;
    const uint64_t leaf_slots = 1ull << table_lg(total_leaves, shift);
    const uint64_t level_slots = (1ull << table_lg(std::max(max_total, nlive), shift)) + kTableOverflow;
    const SortPlan leaf_plan = make_sort_plan(0, 64);

    const uint64_t stage_tiles =
        std::max({tiles(level_slots, kExtractTileItems), tiles(nlive, kBuildTileItems), uint64_t(1)});
    Carver cv;
    const uint64_t o_zero = cv.off;
    const uint64_t o_counters = cv.take(sizeof(BuildCounters));
    const uint64_t o_leaf_sort_zero = cv.take(4 * sort_zero_words(leaf_plan));
    const uint64_t zero_end = cv.off;
    const uint64_t o_sort_counts = cv.take(4 * sort_count_words(total_leaves));
    const uint64_t o_tile_counts = cv.take(4 * stage_tiles);
    const uint64_t o_tile_offsets = cv.take(8 * stage_tiles);
    const uint64_t o_sources = cv.take(sizeof(MergeSource) * nlive * (lc_levels + 1));
    const uint64_t o_leaf_slot_of = cv.take(4 * total_leaves);
    const uint64_t o_leaf_table = cv.take(8 * leaf_slots);
    const uint64_t o_leaf_rank = cv.take(4 * leaf_slots);
    uint64_t o_lk[2], o_ls[2];
    for (int i = 0; i < 2; ++i) {
        o_lk[i] = cv.take(8 * total_leaves);
        o_ls[i] = cv.take(4 * total_leaves);
    }
    const uint64_t o_sorted_slots = cv.take(4 * total_leaves);
    uint64_t o_slot_of[2];
    for (int i = 0; i < 2; ++i) o_slot_of[i] = cv.take(4 * std::max(max_total, nlive));
    const uint64_t o_masks = cv.take(max_total);
    const uint64_t o_children = cv.take(32 * max_total);
    uint64_t o_tk[2], o_trep[2], o_trank[2];
    for (int i = 0; i < 2; ++i) {
        o_tk[i] = cv.take(8 * level_slots);
        o_trep[i] = cv.take(4 * level_slots);
        o_trank[i] = cv.take(4 * level_slots);
    }
    uint64_t o_codes[2];
    for (int i = 0; i < 2; ++i) o_codes[i] = cv.take(8 * nlive);
    TRY(ensure_arena(ctx, cv.off));
    char* A = ctx->arena;
    BuildCounters* c = reinterpret_cast<BuildCounters*>(A + o_counters);
    uint32_t* tile_counts = reinterpret_cast<uint32_t*>(A + o_tile_counts);
    uint64_t* tile_offsets = reinterpret_cast<uint64_t*>(A + o_tile_offsets);

    uint64_t caps[SVDAG_MAX_LEVELS];
    for (int l = 0; l < levels; ++l) caps[l] = l < k ? pow8_capped(l, nlive) : totals[l - k];
    const DagStorage S = plan_storage(levels, caps, total_leaves);
    void* storage = nullptr;
    TRY(device_alloc(ctx, &storage, S.total));
    char* O = static_cast<char*>(storage);
    auto fail = [&](int code) {
        device_release(ctx, storage);
        return code;
    };

    // Source tables: block 0 = leaves, block 1 + l = level l of every live input.
    std::vector<MergeSource> src(nlive * (lc_levels + 1));
    for (int blk = 0; blk <= lc_levels; ++blk) {
        uint64_t first = 0, child_first = 0;
        for (uint64_t j = 0; j < nlive; ++j) {
            const svdag_dag& d = dags[live[j]];
            MergeSource& m = src[uint64_t(blk) * nlive + j];
            if (blk == 0) {
                m = MergeSource{nullptr, nullptr, d.leaves, first, d.num_leaves, 0};
            } else {
                const int l = blk - 1;
                m = MergeSource{d.levels[l].masks, d.levels[l].children, d.levels[l].hashes, first,
                                d.levels[l].num_nodes, child_first};
                child_first += l + 1 < lc_levels ? d.levels[l + 1].num_nodes : d.num_leaves;
            }
            first += m.num;
        }
    }
    MergeSource* d_src = reinterpret_cast<MergeSource*>(A + o_sources);
    if (hipMemcpyAsync(d_src, src.data(), sizeof(MergeSource) * src.size(), hipMemcpyHostToDevice, st) != hipSuccess)
        return fail(set_error(SVDAG_ERR_HIP, "merge source upload failed"));
    std::vector<uint64_t> codes(nlive);
    for (uint64_t j = 0; j < nlive; ++j) codes[j] = live[j];
    if (hipMemcpyAsync(A + o_codes[0], codes.data(), 8 * nlive, hipMemcpyHostToDevice, st) != hipSuccess)
        return fail(set_error(SVDAG_ERR_HIP, "chunk code upload failed"));
    // The upload buffers must outlive the copies: finish them before returning.
    if (hipStreamSynchronize(st) != hipSuccess) return fail(set_error(SVDAG_ERR_HIP, "merge upload failed"));

    launch_fill_u64(st, reinterpret_cast<uint64_t*>(A + o_zero), (zero_end - o_zero) / 8, 0);
    launch_fill_u64(st, &c->leaf_positions, 1, total_leaves);
    for (int l = k; l < levels; ++l) launch_fill_u64(st, &c->pre[l], 1, totals[l - k]);
    if (k > 0) {
        uint64_t groups = 0;
        for (uint64_t j = 0; j < nlive; ++j)
            if (j == 0 || (live[j] >> 3) != (live[j - 1] >> 3)) ++groups;
        launch_fill_u64(st, &c->pre[k - 1], 1, groups);
    }

    // Leaves: shared hashed table (first inserters append to the unique list),
    // sort, rank.
    uint64_t* leaf_table = reinterpret_cast<uint64_t*>(A + o_leaf_table);
    uint32_t* leaf_rank = reinterpret_cast<uint32_t*>(A + o_leaf_rank);
    uint32_t* leaf_slot_of = reinterpret_cast<uint32_t*>(A + o_leaf_slot_of);
    uint32_t* leaf_lg = &c->table_log2[SVDAG_MAX_LEVELS];
    uint64_t* lk0 = reinterpret_cast<uint64_t*>(A + o_lk[0]);
    uint64_t* lk1 = reinterpret_cast<uint64_t*>(A + o_lk[1]);
    uint32_t* ls0 = reinterpret_cast<uint32_t*>(A + o_ls[0]);
    uint32_t* ls1 = reinterpret_cast<uint32_t*>(A + o_ls[1]);
    uint32_t* sorted_slots = reinterpret_cast<uint32_t*>(A + o_sorted_slots);
    launch_table_clear(st, leaf_table, &c->leaf_positions, leaf_lg, shift, 0, leaf_slots);
    launch_merge_insert(st, d_src, uint32_t(nlive), total_leaves, nullptr, nullptr, nullptr, nullptr, leaf_table,
                        nullptr, leaf_lg, leaf_slot_of, c, 0, lk0, ls0);
    uint64_t* leaves_out = reinterpret_cast<uint64_t*>(O + S.leaves);
    if (launch_sort<uint64_t>(st, leaf_plan, lk0, ls0, lk1, lk0, ls1, ls0, leaves_out, sorted_slots, &c->unique_leaves,
                              total_leaves, reinterpret_cast<uint32_t*>(A + o_leaf_sort_zero),
                              reinterpret_cast<uint32_t*>(A + o_sort_counts), nullptr) < 0)
        return fail(set_error(SVDAG_ERR_INTERNAL, "leaf sort plan failed"));
    launch_rank_of_slot(st, sorted_slots, &c->unique_leaves, leaf_rank, total_leaves);

    const uint32_t* child_slot_of = leaf_slot_of;
    const uint32_t* child_rank = leaf_rank;
    uint8_t* masks = reinterpret_cast<uint8_t*>(A + o_masks);
    uint32_t* children = reinterpret_cast<uint32_t*>(A + o_children);
    auto extract = [&](int l, int cur) {
        LevelExtractArgs e{};
        e.table_keys = reinterpret_cast<const uint64_t*>(A + o_tk[cur]);
        e.table_rep = reinterpret_cast<const uint32_t*>(A + o_trep[cur]);
        e.table_log2 = &c->table_log2[l];
        e.rank_of_slot = reinterpret_cast<uint32_t*>(A + o_trank[cur]);
        e.parent_masks = masks;
        e.parent_children = children;
        e.out_hashes = reinterpret_cast<uint64_t*>(O + S.hashes[l]);
        e.out_masks = reinterpret_cast<uint8_t*>(O + S.masks[l]);
        e.out_children = reinterpret_cast<uint32_t*>(O + S.children[l]);
        e.tile_counts = tile_counts;
        e.offsets = tile_offsets;
        e.slots_scratch = &c->table_slots;
        e.q_count = &c->uniq[l];
        e.w_count = &c->words[l];
        launch_level_extract(st, e, level_slots);
    };
    // Levels below the sub-DAG roots.
    int step = 0;
    for (int lc = lc_levels - 1; lc >= 0; --lc, ++step) {
        const int l = lc + k;
        const int cur = step & 1;
        uint64_t* tk = reinterpret_cast<uint64_t*>(A + o_tk[cur]);
        uint32_t* slot_of = reinterpret_cast<uint32_t*>(A + o_slot_of[cur]);
        launch_table_clear(st, tk, &c->pre[l], &c->table_log2[l], shift, kTableOverflow, level_slots);
        launch_merge_insert(st, d_src + uint64_t(lc + 1) * nlive, uint32_t(nlive), totals[lc], child_slot_of,
                            child_rank, masks, children, tk, reinterpret_cast<uint32_t*>(A + o_trep[cur]),
                            &c->table_log2[l], slot_of, c, 1, nullptr, nullptr);
        extract(l, cur);
        child_slot_of = slot_of;
        child_rank = reinterpret_cast<const uint32_t*>(A + o_trank[cur]);
    }
    // New top levels over the chunk roots (DAGCompression.cu:1505-1533).
    for (int l = k - 1; l >= 0; --l, ++step) {
        const int cur = step & 1;
        LevelBuildArgs a{};
        a.child_codes = reinterpret_cast<const uint64_t*>(A + o_codes[(k - 1 - l) & 1]);
        a.child_slots = child_slot_of;
        a.child_rank_of_slot = child_rank;
        a.child_hashes = reinterpret_cast<const uint64_t*>(O + S.hashes[l + 1]);
        a.child_count = &c->pre[l + 1];
        a.parent_codes = reinterpret_cast<uint64_t*>(A + o_codes[(k - l) & 1]);
        a.parent_slots = reinterpret_cast<uint32_t*>(A + o_slot_of[cur]);
        a.parent_masks = masks;
        a.parent_children = children;
        a.table_keys = reinterpret_cast<uint64_t*>(A + o_tk[cur]);
        a.table_rep = reinterpret_cast<uint32_t*>(A + o_trep[cur]);
        a.table_log2 = &c->table_log2[l];
        a.tile_counts = tile_counts;
        a.offsets = tile_offsets;
        a.next_count = l > 0 ? &c->pre[l - 1] : nullptr;
        a.counters = c;
        launch_table_clear(st, a.table_keys, &c->pre[l], &c->table_log2[l], shift, kTableOverflow, level_slots);
        launch_level_build(st, a, nlive);
        extract(l, cur);
        child_slot_of = a.parent_slots;
        child_rank = reinterpret_cast<const uint32_t*>(A + o_trank[cur]);
    }
    {
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return fail(set_error(SVDAG_ERR_HIP, "kernel launch failed: %s", hipGetErrorString(e)));
    }
    if (hipMemcpyAsync(ctx->h_counters, c, sizeof(BuildCounters), hipMemcpyDeviceToHost, st) != hipSuccess)
        return fail(set_error(SVDAG_ERR_HIP, "counter read-back failed: %s", hipGetErrorString(hipGetLastError())));
    mark(ctx, "counter read-back");
    if (const int rc = finish(ctx, "merge_dags")) return fail(rc);
    const BuildCounters& h = *ctx->h_counters;
    if (h.error & kErrTableOverflow) return fail(set_error(SVDAG_ERR_INTERNAL, "dedup table probe overflow"));
    if (h.error & kErrInternal) return fail(set_error(SVDAG_ERR_INTERNAL, "leaf reduce region overflow"));
    if (h.uniq[0] != 1) return fail(set_error(SVDAG_ERR_INTERNAL, "merged root level holds %llu nodes", (unsigned long long)h.uniq[0]));
    out->storage = storage;
    out->leaves = leaves_out;
    out->num_leaves = h.unique_leaves;
    for (int l = 0; l < levels; ++l) {
        out->levels[l].hashes = reinterpret_cast<uint64_t*>(O + S.hashes[l]);
        out->levels[l].children = reinterpret_cast<uint32_t*>(O + S.children[l]);
        out->levels[l].masks = reinterpret_cast<uint8_t*>(O + S.masks[l]);
        out->levels[l].num_nodes = h.uniq[l];
        out->level_words[l] = h.words[l];
    }
    return 0;
}

// SortKeys (src/DAGCompression.cu:336-357: cub::DeviceRadixSort::SortKeys on
// bits [0, bits)), stable like CUB's: every pass ranks stably.
template <typename K>
int sort_keys_entry(svdag_context* ctx, const K* d_keys, uint64_t n, int32_t bits, K* d_out)
{
    TRY(check_ctx(ctx));
    const int key_bits = int(8 * sizeof(K));
    if (bits < 1 || bits > key_bits) return set_error(SVDAG_ERR_ARGUMENT, "bits %d outside [1, %d]", bits, key_bits);
    if (n >= (1ull << 32)) return set_error(SVDAG_ERR_ARGUMENT, "n = %llu keys: at most 2^32 - 1", (unsigned long long)n);
    if (n && (!d_keys || !d_out)) return set_error(SVDAG_ERR_ARGUMENT, "null keys or output");
    if (n && d_keys == d_out) return set_error(SVDAG_ERR_ARGUMENT, "output must not alias the input");
    if (n == 0) return 0;
    hipStream_t st = ctx->stream;
    SortPlan plan = make_sort_plan(0, bits, ctx->radix_bits, ctx->sort_ipt);
    if (plan.passes == 0) return set_error(SVDAG_ERR_INTERNAL, "sort plan failed");
    plan.stable = true;
    Carver cv;
    const uint64_t o_zero = cv.take(4 * sort_zero_words(plan));
    const uint64_t o_n = cv.take(8), o_or = cv.take(8);
    const uint64_t o_counts = cv.take(4 * sort_count_words(n));
    const uint64_t o_a = cv.take(sizeof(K) * n), o_b = cv.take(sizeof(K) * n);
    TRY(ensure_arena(ctx, cv.off));
    char* A = ctx->arena;
    launch_fill_u64(st, reinterpret_cast<uint64_t*>(A + o_zero), (o_n - o_zero) / 8, 0);
    launch_fill_u64(st, reinterpret_cast<uint64_t*>(A + o_n), 1, n);
    launch_fill_u64(st, reinterpret_cast<uint64_t*>(A + o_or), 1, 0);
    const int where = launch_sort<K>(st, plan, d_keys, nullptr, reinterpret_cast<K*>(A + o_a),
                                     reinterpret_cast<K*>(A + o_b), nullptr, nullptr, d_out, nullptr,
                                     reinterpret_cast<const uint64_t*>(A + o_n), n,
                                     reinterpret_cast<uint32_t*>(A + o_zero), reinterpret_cast<uint32_t*>(A + o_counts),
                                     reinterpret_cast<unsigned long long*>(A + o_or));
    if (where != 2) return set_error(SVDAG_ERR_INTERNAL, "sort plan failed");
    HIP_TRY(hipGetLastError());
    return finish(ctx, "sort_keys");
}

}  // namespace

extern "C" {

int svdag_context_create(svdag_context** out, void* stream)
{
    if (!out) return set_error(SVDAG_ERR_ARGUMENT, "null output");
    auto* ctx = new svdag_context();
    ctx->stream = static_cast<hipStream_t>(stream);
    if (hipGetDevice(&ctx->device) != hipSuccess) {
        delete ctx;
        return set_error(SVDAG_ERR_HIP, "no HIP device");
    }
    if (hipHostMalloc(reinterpret_cast<void**>(&ctx->h_counters), sizeof(BuildCounters), hipHostMallocDefault) !=
        hipSuccess) {
        delete ctx;
        return set_error(SVDAG_ERR_HIP, "cannot allocate pinned staging");
    }
    for (int i = 0; i < svdag_context::kMaxMarks; ++i)
        if (hipEventCreateWithFlags(&ctx->marks[i], hipEventDisableTiming) != hipSuccess) ctx->marks[i] = nullptr;
    (void)hipGetLastError();
    *out = ctx;
    return 0;
}

int svdag_context_destroy(svdag_context* ctx)
{
    if (!ctx) return 0;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->arena) (void)hipFree(ctx->arena);
    for (auto& kv : ctx->block_size) (void)hipFree(kv.first);  // DAGs still alive become invalid
    if (ctx->h_counters) (void)hipHostFree(ctx->h_counters);
    for (auto& e : ctx->marks)
        if (e) (void)hipEventDestroy(e);
    if (ctx->side) {
        (void)hipStreamSynchronize(ctx->side);
        (void)hipEventDestroy(ctx->fork_ev);
        (void)hipEventDestroy(ctx->join_ev);
        (void)hipStreamDestroy(ctx->side);
    }
    if (active_recorder() == &ctx->recorder) active_recorder() = nullptr;
    delete ctx;
    return 0;
}

int svdag_set_option(svdag_context* ctx, int32_t option, int64_t value)
{
    if (!ctx) return set_error(SVDAG_ERR_ARGUMENT, "null context");
    switch (option) {
    case SVDAG_OPT_CHECK_COLLISIONS: ctx->check_collisions = value != 0; return 0;
    case SVDAG_OPT_MEMORY_BUDGET: ctx->budget = uint64_t(std::max<int64_t>(0, value)); return 0;
    case SVDAG_OPT_TABLE_LOAD_SHIFT:
        if (value < 1 || value > 3) return set_error(SVDAG_ERR_ARGUMENT, "load shift must be 1..3");
        ctx->load_shift = uint32_t(value);
        return 0;
    case SVDAG_OPT_HOST_THREADS: ctx->host_threads = int(std::max<int64_t>(1, value)); return 0;
    case SVDAG_OPT_WAIT_SECONDS: ctx->wait_seconds = double(std::max<int64_t>(1, value)); return 0;
    case SVDAG_OPT_RADIX_BITS:
        if (value < 4 || value > kMaxRadixBits) return set_error(SVDAG_ERR_ARGUMENT, "radix bits must be 4..%d", kMaxRadixBits);
        ctx->radix_bits = int(value);
        return 0;
    case SVDAG_OPT_TABLE_ESTIMATE: ctx->table_estimate = value != 0; return 0;
    case SVDAG_OPT_SEG_FIRST_BITS:
        if (value < 0 || value > kMaxSegDigitBits)
            return set_error(SVDAG_ERR_ARGUMENT, "segmented first-pass bits must be 0..%u", kMaxSegDigitBits);
        ctx->seg_first_bits = uint32_t(value);
        return 0;
    case SVDAG_OPT_VOXEL_STAGE_BYTES:
        if (value < 0) return set_error(SVDAG_ERR_ARGUMENT, "stage bytes must be >= 0");
        ctx->vox_stage_max = uint64_t(value);
        return 0;
    case SVDAG_OPT_SORT_TILE:
        if (value != 4096 && value != 8192) return set_error(SVDAG_ERR_ARGUMENT, "sort tile must be 4096 or 8192 keys");
        ctx->sort_ipt = value == 8192 ? kSortIptLarge : kSortIptSmall;
        return 0;
    case SVDAG_OPT_TOP_DIGIT_FIRST:
        if (value < kTopFirstNever || value > kTopFirstAlways)
            return set_error(SVDAG_ERR_ARGUMENT, "top digit first must be 0 (never), 1 (auto) or 2 (always)");
        ctx->top_first = int(value);
        return 0;
    default: return set_error(SVDAG_ERR_ARGUMENT, "unknown option %d", option);
    }
}

uint64_t svdag_scratch_bytes(uint64_t n, int32_t depth, int32_t key_bytes)
{
    if (depth < 3 || depth > 21 || n == 0) return 0;
    // The largest layout over the digit widths (narrow digits: more passes and
    // histograms; wide ones: the top-digit-first tables).
    uint64_t most = 0;
    for (int bits = 4; bits <= kMaxRadixBits; ++bits)
        most = std::max(most, plan_build(n, depth, key_bytes == 4 ? 4 : 8, 1, false, bits, kSortIptSmall,
                                         kTopFirstAlways).total);
    return most;
}

int svdag_create_dag(svdag_context* ctx, const uint64_t* d_keys, uint64_t n, int32_t depth, svdag_dag* out)
{
    TRY(check_ctx(ctx));
    if (!out) return set_error(SVDAG_ERR_ARGUMENT, "null output");
    if (depth < 3 || depth > 21) return set_error(SVDAG_ERR_ARGUMENT, "depth %d outside [3, 21]", depth);
    if (n >= (1ull << 30)) return set_error(SVDAG_ERR_ARGUMENT, "%llu keys: at most 2^30 - 1 per call", (unsigned long long)n);
    if (n && !d_keys) return set_error(SVDAG_ERR_ARGUMENT, "null keys");
    const int rc = build<uint64_t>(ctx, d_keys, n, depth, out, nullptr, ctx->table_estimate);
    return rc == kRetryWithBounds ? build<uint64_t>(ctx, d_keys, n, depth, out, nullptr, false) : rc;
}

}  // extern "C"

namespace {

// Device address of `scratch` on each device, resolved once per device (with
// that device current) on the first no-allocation call there.
constexpr int kMaxDevices = 64;
std::mutex g_fixed_mutex;
char* g_fixed_scratch[kMaxDevices] = {};

int fixed_scratch(int device, char** out)
{
    if (device < 0 || device >= kMaxDevices) return set_error(SVDAG_ERR_ARGUMENT, "device %d out of range", device);
    std::lock_guard<std::mutex> lock(g_fixed_mutex);
    if (!g_fixed_scratch[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(SVDAG_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device,
                             hipGetErrorString(e));
        g_fixed_scratch[device] = static_cast<char*>(p);
    }
    *out = g_fixed_scratch[device];
    return 0;
}

}  // namespace

extern "C" {

uint64_t svdag_result_bytes(uint64_t n, int32_t depth)
{
    if (depth < 3 || depth > 21 || n == 0) return 0;
    return result_bytes(n, depth);
}

int svdag_create_dag_noalloc(svdag_context* ctx, const uint64_t* d_keys, uint64_t n, int32_t depth,
                             void* d_storage, uint64_t storage_bytes, svdag_dag* out)
{
    TRY(check_ctx(ctx));
    // Kernel timing creates its events on demand: off for this call.
    active_recorder() = nullptr;
    if (!out) return set_error(SVDAG_ERR_ARGUMENT, "null output");
    if (depth < 3 || depth > 21) return set_error(SVDAG_ERR_ARGUMENT, "depth %d outside [3, 21]", depth);
    if (n >= (1ull << 30)) return set_error(SVDAG_ERR_ARGUMENT, "%llu keys: at most 2^30 - 1 per call", (unsigned long long)n);
    if (n && (!d_keys || !d_storage)) return set_error(SVDAG_ERR_ARGUMENT, "null keys or storage");
    char* base = nullptr;
    if (n) TRY(fixed_scratch(ctx->device, &base));
    const FixedMemory fm{base, SVDAG_NOALLOC_SCRATCH_BYTES, static_cast<char*>(d_storage), storage_bytes};
    const int rc = build<uint64_t>(ctx, d_keys, n, depth, out, nullptr, ctx->table_estimate, &fm);
    return rc == kRetryWithBounds ? build<uint64_t>(ctx, d_keys, n, depth, out, nullptr, false, &fm) : rc;
}

int svdag_create_dag_u32(svdag_context* ctx, const uint32_t* d_keys, uint64_t n, int32_t depth, svdag_dag* out)
{
    TRY(check_ctx(ctx));
    if (!out) return set_error(SVDAG_ERR_ARGUMENT, "null output");
    if (depth < 3 || depth > 10) return set_error(SVDAG_ERR_ARGUMENT, "depth %d outside [3, 10] for 32-bit keys", depth);
    if (n >= (1ull << 30)) return set_error(SVDAG_ERR_ARGUMENT, "%llu keys: at most 2^30 - 1 per call", (unsigned long long)n);
    if (n && !d_keys) return set_error(SVDAG_ERR_ARGUMENT, "null keys");
    const int rc = build<uint32_t>(ctx, d_keys, n, depth, out, nullptr, ctx->table_estimate);
    return rc == kRetryWithBounds ? build<uint32_t>(ctx, d_keys, n, depth, out, nullptr, false) : rc;
}

int svdag_create_dag_batched(svdag_context* ctx, const uint64_t* d_keys, const uint64_t* offsets, uint64_t num_chunks,
                             int32_t chunk_depth, svdag_dag* out)
{
    TRY(check_ctx(ctx));
    if (!out || !offsets) return set_error(SVDAG_ERR_ARGUMENT, "null argument");
    int extra = 0;
    for (uint64_t c = num_chunks; c > 1; c /= 8) {
        if (c % 8) return set_error(SVDAG_ERR_ARGUMENT, "chunk count %llu is not a power of 8", (unsigned long long)num_chunks);
        ++extra;
    }
    if (num_chunks == 0) return set_error(SVDAG_ERR_ARGUMENT, "no chunks");
    const int depth = chunk_depth + extra;
    if (chunk_depth < 3 || depth > 21) return set_error(SVDAG_ERR_ARGUMENT, "bad chunk depth %d", chunk_depth);
    for (uint64_t i = 0; i < num_chunks; ++i)
        if (offsets[i + 1] < offsets[i]) return set_error(SVDAG_ERR_ARGUMENT, "offsets decrease at chunk %llu", (unsigned long long)i);
    const uint64_t n = offsets[num_chunks] - offsets[0];
    if (offsets[0] != 0) return set_error(SVDAG_ERR_ARGUMENT, "offsets[0] must be 0");
    if (n >= (1ull << 30)) return set_error(SVDAG_ERR_ARGUMENT, "%llu keys: at most 2^30 - 1 per call", (unsigned long long)n);
    if (n && !d_keys) return set_error(SVDAG_ERR_ARGUMENT, "null keys");
    BatchInfo b{offsets, num_chunks, chunk_depth};
    const int rc = build<uint64_t>(ctx, d_keys, n, depth, out, &b, ctx->table_estimate);
    return rc == kRetryWithBounds ? build<uint64_t>(ctx, d_keys, n, depth, out, &b, false) : rc;
}

int svdag_create_dag_batched_gather(svdag_context* ctx, const uint64_t* const* d_chunk_keys, const uint64_t* counts,
                                    uint64_t num_chunks, int32_t chunk_depth, svdag_dag* out)
{
    TRY(check_ctx(ctx));
    if (!d_chunk_keys || !counts || !out) return set_error(SVDAG_ERR_ARGUMENT, "null argument");
    if (num_chunks == 0 || num_chunks > (1ull << 21)) return set_error(SVDAG_ERR_ARGUMENT, "chunk count out of range");
    std::vector<uint64_t> offsets(num_chunks + 1, 0);
    for (uint64_t i = 0; i < num_chunks; ++i) {
        if (counts[i] && !d_chunk_keys[i]) return set_error(SVDAG_ERR_ARGUMENT, "chunk %llu has keys but no pointer", (unsigned long long)i);
        offsets[i + 1] = offsets[i] + counts[i];
        if (offsets[i + 1] >= (1ull << 30)) return set_error(SVDAG_ERR_ARGUMENT, "more than 2^30 - 1 keys");
    }
    const uint64_t n = offsets[num_chunks];
    // Chunks that already sit back to back in chunk order need no copy.
    uintptr_t base = 0;
    bool have_base = false, contiguous = true;
    for (uint64_t i = 0; i < num_chunks && contiguous; ++i) {
        if (!counts[i]) continue;
        const uintptr_t b = reinterpret_cast<uintptr_t>(d_chunk_keys[i]) - 8 * offsets[i];
        if (!have_base) base = b;
        have_base = true;
        contiguous = b == base;
    }
    if (contiguous)
        return svdag_create_dag_batched(ctx, have_base ? reinterpret_cast<const uint64_t*>(base) : nullptr,
                                        offsets.data(), num_chunks, chunk_depth, out);
    void* keys = nullptr;
    if (n) {
        // The chunks' keys back to back, in chunk order, in a pooled block
        // (released in stream order once the build has read them).
        TRY(device_alloc(ctx, &keys, 8 * n));
        // The host table stays in the context until the next gathered build,
        // by which time this call's work (and so the copy) has finished.
        std::vector<GatherSource>& src = ctx->gather_src;
        src.clear();
        for (uint64_t i = 0; i < num_chunks; ++i)
            if (counts[i]) src.push_back(GatherSource{d_chunk_keys[i], offsets[i]});
        void* table = nullptr;
        int rc = device_alloc(ctx, &table, sizeof(GatherSource) * src.size());
        if (rc == 0 && hipMemcpyAsync(table, src.data(), sizeof(GatherSource) * src.size(), hipMemcpyHostToDevice,
                                      ctx->stream) != hipSuccess)
            rc = set_error(SVDAG_ERR_HIP, "gather table upload failed");
        if (rc == 0) {
            launch_gather_keys(ctx->stream, static_cast<const GatherSource*>(table), uint32_t(src.size()), n,
                               static_cast<uint64_t*>(keys));
            if (hipGetLastError() != hipSuccess) rc = set_error(SVDAG_ERR_HIP, "gather launch failed");
        }
        device_release(ctx, table);
        if (rc != 0) {
            device_release(ctx, keys);
            return rc;
        }
    }
    const int rc = svdag_create_dag_batched(ctx, static_cast<const uint64_t*>(keys), offsets.data(), num_chunks,
                                            chunk_depth, out);
    device_release(ctx, keys);
    return rc;
}

int svdag_merge_dags(svdag_context* ctx, const svdag_dag* dags, uint64_t count, svdag_dag* out)
{
    TRY(check_ctx(ctx));
    if (!dags || !out || count == 0) return set_error(SVDAG_ERR_ARGUMENT, "null argument or no DAG");
    int k = 0;
    for (uint64_t c = count; c > 1; c /= 8) {
        if (c % 8) return set_error(SVDAG_ERR_ARGUMENT, "DAG count %llu is not a power of 8", (unsigned long long)count);
        ++k;
    }
    const int lc_levels = dags[0].num_levels;
    if (lc_levels < 1 || lc_levels + k > SVDAG_MAX_LEVELS || lc_levels + k + 2 > 21)
        return set_error(SVDAG_ERR_ARGUMENT, "merged depth %d out of range", lc_levels + k + 2);
    for (uint64_t i = 0; i < count; ++i) {
        if (dags[i].num_levels != lc_levels) return set_error(SVDAG_ERR_ARGUMENT, "DAG %llu has a different depth", (unsigned long long)i);
        if (dags[i].num_leaves && dags[i].levels[0].num_nodes != 1)  // DAGCompression.cu:1476
            return set_error(SVDAG_ERR_ARGUMENT, "DAG %llu has %llu root nodes", (unsigned long long)i,
                             (unsigned long long)dags[i].levels[0].num_nodes);
    }
    // Device residency for the inputs (host inputs are uploaded temporarily).
    std::vector<svdag_dag> dev(dags, dags + count);
    std::vector<char> owned(count, 0);
    int rc = 0;
    for (uint64_t i = 0; i < count && rc == 0; ++i)
        if (dags[i].location == SVDAG_HOST && dags[i].num_leaves) {
            rc = svdag_dag_to_device(ctx, &dags[i], &dev[i]);
            owned[i] = rc == 0;
        }
    if (rc == 0) rc = merge_device(ctx, dev.data(), count, k, out);
    const std::string err = g_error;
    (void)hipStreamSynchronize(ctx->stream);
    for (uint64_t i = 0; i < count; ++i)
        if (owned[i]) svdag_free_dag(ctx, &dev[i]);
    g_error = err;
    return rc;
}

int svdag_create_final_dag(svdag_context* ctx, const svdag_dag* dag, uint32_t** d_words, uint64_t* n_words)
{
    TRY(check_ctx(ctx));
    if (!dag || !d_words || !n_words) return set_error(SVDAG_ERR_ARGUMENT, "null argument");
    *d_words = nullptr;
    *n_words = 0;
    if (dag->location != SVDAG_DEVICE) return set_error(SVDAG_ERR_ARGUMENT, "create_final_dag needs a device-resident DAG");
    if (dag->num_levels < 1 || dag->num_levels > SVDAG_MAX_LEVELS) return set_error(SVDAG_ERR_ARGUMENT, "bad level count");
    const uint64_t total_words = dag_words(dag);
    if (total_words > 0xFFFFFFFFull)
        return set_error(SVDAG_ERR_CAPACITY, "final DAG of %llu words does not fit uint32 offsets", (unsigned long long)total_words);
    FinalArgs a{};
    a.num_levels = dag->num_levels;
    uint64_t nodes = 0;
    for (int l = 0; l < dag->num_levels; ++l) {
        a.levels[l].masks = dag->levels[l].masks;
        a.levels[l].children = dag->levels[l].children;
        a.levels[l].first_node = nodes;
        a.levels[l].num_nodes = dag->levels[l].num_nodes;
        nodes += dag->levels[l].num_nodes;
    }
    a.total_nodes = nodes;
    a.leaves = dag->leaves;
    a.num_leaves = dag->num_leaves;
    a.leaves_start = total_words - 2 * dag->num_leaves;
    const uint64_t ntiles = std::max<uint64_t>(1, tiles(nodes, kFinalTileItems));
    Carver cv;
    const uint64_t o_counts = cv.take(4 * ntiles);
    const uint64_t o_offsets = cv.take(8 * ntiles);
    const uint64_t o_pos = cv.take(4 * std::max<uint64_t>(nodes, 1));
    TRY(ensure_arena(ctx, cv.off));
    void* words = nullptr;
    TRY(device_alloc(ctx, &words, 4 * std::max<uint64_t>(total_words, 1)));
    a.words = static_cast<uint32_t*>(words);
    a.positions = reinterpret_cast<uint32_t*>(ctx->arena + o_pos);
    a.counts = reinterpret_cast<uint32_t*>(ctx->arena + o_counts);
    a.offsets = reinterpret_cast<uint64_t*>(ctx->arena + o_offsets);
    launch_final(ctx->stream, a);
    const hipError_t e = hipGetLastError();
    mark(ctx, "final layout");
    if (e != hipSuccess) {
        device_release(ctx, words);
        return set_error(SVDAG_ERR_HIP, "final layout failed: %s", hipGetErrorString(e));
    }
    if (const int rc = finish(ctx, "create_final_dag")) {
        device_release(ctx, words);
        return rc;
    }
    *d_words = static_cast<uint32_t*>(words);
    *n_words = total_words;
    return 0;
}

int svdag_dag_to_host(svdag_context* ctx, const svdag_dag* dag, svdag_dag* out)
{
    TRY(check_ctx(ctx));
    if (!dag || !out) return set_error(SVDAG_ERR_ARGUMENT, "null argument");
    uint64_t total = 8 * dag->num_leaves + 8;
    for (int l = 0; l < dag->num_levels; ++l) total += 41 * dag->levels[l].num_nodes + 32;
    char* mem = static_cast<char*>(host_alloc(total));
    if (!mem) return set_error(SVDAG_ERR_MEMORY, "host allocation failed");
    *out = *dag;
    out->location = SVDAG_HOST;
    out->storage = mem;
    const hipMemcpyKind kind = dag->location == SVDAG_DEVICE ? hipMemcpyDeviceToHost : hipMemcpyHostToHost;
    char* p = mem;
    auto copy = [&](void* src, uint64_t bytes) -> void* {
        void* dst = p;
        if (bytes) (void)hipMemcpyAsync(dst, src, bytes, kind, ctx->stream);
        p += (bytes + 7) / 8 * 8;
        return dst;
    };
    for (int l = 0; l < dag->num_levels; ++l) {
        const uint64_t q = dag->levels[l].num_nodes;
        out->levels[l].hashes = static_cast<uint64_t*>(copy(dag->levels[l].hashes, 8 * q));
        out->levels[l].children = static_cast<uint32_t*>(copy(dag->levels[l].children, 32 * q));
        out->levels[l].masks = static_cast<uint8_t*>(copy(dag->levels[l].masks, q));
    }
    out->leaves = static_cast<uint64_t*>(copy(dag->leaves, 8 * dag->num_leaves));
    const hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        host_release(mem);
        out->storage = nullptr;
        return set_error(SVDAG_ERR_HIP, "copy to host failed: %s", hipGetErrorString(e));
    }
    return 0;
}

int svdag_dag_to_device(svdag_context* ctx, const svdag_dag* dag, svdag_dag* out)
{
    TRY(check_ctx(ctx));
    if (!dag || !out) return set_error(SVDAG_ERR_ARGUMENT, "null argument");
    if (dag->num_levels < 1 || dag->num_levels > SVDAG_MAX_LEVELS) return set_error(SVDAG_ERR_ARGUMENT, "bad level count");
    if (dag->location == SVDAG_HOST) {
        // Host DAGs come from callers: check every child index before any kernel reads through it.
        for (int l = 0; l < dag->num_levels; ++l) {
            const svdag_level& in = dag->levels[l];
            const uint64_t below = l + 1 < dag->num_levels ? dag->levels[l + 1].num_nodes : dag->num_leaves;
            for (uint64_t i = 0; i < in.num_nodes; ++i)
                for (int s = 0; s < 8; ++s) {
                    const uint32_t c = in.children[8 * i + s];
                    const bool set = (in.masks[i] >> s) & 1;
                    if (set != (c != kEmptyChild) || (set && c >= below))
                        return set_error(SVDAG_ERR_ARGUMENT, "level %d node %llu slot %d: child %u inconsistent", l,
                                         (unsigned long long)i, s, c);
                }
        }
    }
    uint64_t caps[SVDAG_MAX_LEVELS];
    for (int l = 0; l < dag->num_levels; ++l) caps[l] = dag->levels[l].num_nodes;
    const DagStorage S = plan_storage(dag->num_levels, caps, dag->num_leaves);
    void* storage = nullptr;
    TRY(device_alloc(ctx, &storage, S.total));
    char* O = static_cast<char*>(storage);
    *out = *dag;
    out->location = SVDAG_DEVICE;
    out->storage = storage;
    const hipMemcpyKind kind = dag->location == SVDAG_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    for (int l = 0; l < dag->num_levels; ++l) {
        const svdag_level& in = dag->levels[l];
        const uint64_t q = in.num_nodes;
        out->levels[l].hashes = reinterpret_cast<uint64_t*>(O + S.hashes[l]);
        out->levels[l].children = reinterpret_cast<uint32_t*>(O + S.children[l]);
        out->levels[l].masks = reinterpret_cast<uint8_t*>(O + S.masks[l]);
        if (q) {
            (void)hipMemcpyAsync(out->levels[l].hashes, in.hashes, 8 * q, kind, ctx->stream);
            (void)hipMemcpyAsync(out->levels[l].children, in.children, 32 * q, kind, ctx->stream);
            (void)hipMemcpyAsync(out->levels[l].masks, in.masks, q, kind, ctx->stream);
        }
        if (dag->location == SVDAG_HOST) {  // level words from the masks
            uint64_t w = 0;
            for (uint64_t i = 0; i < q; ++i) w += 1 + __builtin_popcount(in.masks[i]);
            out->level_words[l] = w;
        }
    }
    out->leaves = reinterpret_cast<uint64_t*>(O + S.leaves);
    if (dag->num_leaves) (void)hipMemcpyAsync(out->leaves, dag->leaves, 8 * dag->num_leaves, kind, ctx->stream);
    const hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        device_release(ctx, storage);
        out->storage = nullptr;
        return set_error(SVDAG_ERR_HIP, "copy to device failed: %s", hipGetErrorString(e));
    }
    return 0;
}

int svdag_free_dag(svdag_context* ctx, svdag_dag* dag)
{
    if (!dag) return 0;
    if (dag->storage) {
        if (dag->location == SVDAG_DEVICE) {
            if (!ctx) return set_error(SVDAG_ERR_ARGUMENT, "device DAG needs its context to be freed");
            (void)hipSetDevice(ctx->device);
            device_release(ctx, dag->storage);
        } else {
            host_release(dag->storage);
        }
    }
    std::memset(dag, 0, sizeof(*dag));
    return 0;
}

int svdag_free_words(svdag_context* ctx, uint32_t* d_words)
{
    TRY(check_ctx(ctx));
    device_release(ctx, d_words);
    return 0;
}

int svdag_memory_usage(svdag_context* ctx, uint64_t* scratch_bytes, uint64_t* result_bytes, uint64_t* cached_bytes)
{
    if (!ctx) return set_error(SVDAG_ERR_ARGUMENT, "null context");
    uint64_t cached = 0;
    for (const auto& kv : ctx->free_blocks) cached += kv.first;
    if (scratch_bytes) *scratch_bytes = ctx->arena_cap;
    if (result_bytes) *result_bytes = ctx->pool_bytes - cached;
    if (cached_bytes) *cached_bytes = cached;
    return 0;
}

int svdag_free_all(svdag_context* ctx)
{
    TRY(check_ctx(ctx));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (ctx->arena) HIP_TRY(hipFree(ctx->arena));
    ctx->arena = nullptr;
    ctx->arena_cap = 0;
    pool_trim(ctx);
    host_trim();
    return 0;
}

const char* svdag_last_error(void) { return g_error.c_str(); }

// The reference's DAG file (FFileWriter / FFileReader framing, src/Serializer.h:10-101,
// read in that order at src/DAG_Compression_Fast.cpp:362-375): six float64
// (AABB min x y z, max x y z), uint32 level count, then the words as an array
// (uint64 count, uint32 data), native little-endian, no padding.
int svdag_write_dag_file(svdag_context* ctx, const char* path, const double* aabb_min, const double* aabb_max,
                         uint32_t levels, const uint32_t* words, uint64_t n_words, int32_t words_location)
{
    if (!path || !aabb_min || !aabb_max || (n_words && !words))
        return set_error(SVDAG_ERR_ARGUMENT, "null path, bounds or words");
    if (words_location != SVDAG_HOST && words_location != SVDAG_DEVICE)
        return set_error(SVDAG_ERR_ARGUMENT, "bad words location %d", words_location);
    std::vector<uint32_t> staged;
    const uint32_t* src = words;
    if (words_location == SVDAG_DEVICE && n_words) {
        TRY(check_ctx(ctx));
        staged.resize(n_words);
        HIP_TRY(hipMemcpyAsync(staged.data(), words, 4 * n_words, hipMemcpyDeviceToHost, ctx->stream));
        TRY(finish(ctx, "write_dag_file"));
        src = staged.data();
    }
    FILE* f = std::fopen(path, "wb");
    if (!f) return set_error(SVDAG_ERR_ARGUMENT, "cannot open %s for writing", path);
    bool ok = std::fwrite(aabb_min, sizeof(double), 3, f) == 3 && std::fwrite(aabb_max, sizeof(double), 3, f) == 3 &&
              std::fwrite(&levels, sizeof(uint32_t), 1, f) == 1 && std::fwrite(&n_words, sizeof(uint64_t), 1, f) == 1 &&
              (n_words == 0 || std::fwrite(src, sizeof(uint32_t), n_words, f) == n_words);
    ok = std::fclose(f) == 0 && ok;
    if (!ok) return set_error(SVDAG_ERR_ARGUMENT, "write to %s failed", path);
    return 0;
}

int svdag_read_dag_file(const char* path, double* aabb_min, double* aabb_max, uint32_t* levels, uint32_t** words,
                        uint64_t* n_words)
{
    if (!path || !aabb_min || !aabb_max || !levels || !words || !n_words)
        return set_error(SVDAG_ERR_ARGUMENT, "null argument");
    *words = nullptr;
    *n_words = 0;
    FILE* f = std::fopen(path, "rb");
    if (!f) return set_error(SVDAG_ERR_ARGUMENT, "cannot open %s", path);
    uint64_t n = 0;
    bool ok = std::fread(aabb_min, sizeof(double), 3, f) == 3 && std::fread(aabb_max, sizeof(double), 3, f) == 3 &&
              std::fread(levels, sizeof(uint32_t), 1, f) == 1 && std::fread(&n, sizeof(uint64_t), 1, f) == 1;
    // The word count must fit what is left of the file.
    long here = ok ? std::ftell(f) : -1;
    if (ok && here >= 0 && std::fseek(f, 0, SEEK_END) == 0) {
        const long end = std::ftell(f);
        ok = end >= here && n <= uint64_t(end - here) / sizeof(uint32_t) && std::fseek(f, here, SEEK_SET) == 0;
    } else {
        ok = false;
    }
    uint32_t* buf = nullptr;
    if (ok && n) {
        buf = static_cast<uint32_t*>(std::malloc(sizeof(uint32_t) * n));
        ok = buf && std::fread(buf, sizeof(uint32_t), n, f) == n;
    }
    std::fclose(f);
    if (!ok) {
        std::free(buf);
        return set_error(SVDAG_ERR_ARGUMENT, "%s is not a DAG file (short or inconsistent)", path);
    }
    *words = buf;
    *n_words = n;
    return 0;
}

int svdag_voxelize_triangles(svdag_context* ctx, const double* d_tris, uint64_t n_tris, int32_t depth,
                             uint64_t** d_keys, uint64_t* n_keys)
{
    TRY(check_ctx(ctx));
    if (!d_keys || !n_keys || (n_tris && !d_tris)) return set_error(SVDAG_ERR_ARGUMENT, "null argument");
    if (depth < 1 || depth > 21) return set_error(SVDAG_ERR_ARGUMENT, "depth %d outside [1, 21]", depth);
    *d_keys = nullptr;
    *n_keys = 0;
    if (n_tris == 0) return 0;
    hipStream_t st = ctx->stream;
    // Triangle stage in the arena: constants, row counts and offsets, and the
    // counters [triangles, rows, columns, keys, rectangle columns, error].
    Carver cv;
    const uint64_t o_setup = cv.take(sizeof(TriSetup) * n_tris);
    const uint64_t o_rows = cv.take(4 * n_tris);
    const uint64_t o_roff = cv.take(8 * n_tris);
    const uint64_t o_cnt = cv.take(8 * 6);
    TRY(ensure_arena(ctx, cv.off));
    char* A = ctx->arena;
    TriSetup* setup = reinterpret_cast<TriSetup*>(A + o_setup);
    uint32_t* rows = reinterpret_cast<uint32_t*>(A + o_rows);
    uint64_t* roff = reinterpret_cast<uint64_t*>(A + o_roff);
    uint64_t* cnt = reinterpret_cast<uint64_t*>(A + o_cnt);
    launch_fill_u64(st, cnt, 6, 0);
    launch_fill_u64(st, cnt, 1, n_tris);
    launch_tri_setup(st, d_tris, n_tris, depth, setup, rows, cnt + 4, reinterpret_cast<unsigned long long*>(cnt + 5));
    launch_scan_counts(st, rows, roff, cnt, 1, cnt + 1);
    uint64_t h[6] = {};
    HIP_TRY(hipMemcpyAsync(h, cnt, sizeof(h), hipMemcpyDeviceToHost, st));
    TRY(finish(ctx, "voxelize (triangle setup)"));
    if (h[5] & kErrKeyRange) return set_error(SVDAG_ERR_ARGUMENT, "a triangle has a non-finite coordinate");
    const uint64_t n_rows = h[1], rect_columns = h[4];
    if (n_rows == 0 || rect_columns == 0) return 0;
    // Row table and tile arrays (sized for the rectangles' columns, which bound
    // the kept ones), then the keys (the caller's, svdag_free_keys).
    const uint64_t ntiles = tiles(rect_columns, kVoxTileItems);
    Carver rows_cv;
    const uint64_t o_tri = rows_cv.take(4 * n_rows), o_u = rows_cv.take(4 * n_rows), o_w0 = rows_cv.take(4 * n_rows);
    const uint64_t o_cols = rows_cv.take(4 * n_rows), o_coff = rows_cv.take(8 * n_rows);
    const uint64_t o_tcnt = rows_cv.take(4 * ntiles), o_toff = rows_cv.take(8 * ntiles), o_trow = rows_cv.take(8 * (ntiles + 1));
    const uint64_t nrt = tiles(n_rows, kVoxTileItems);
    const uint64_t o_rtc = rows_cv.take(4 * nrt), o_rto = rows_cv.take(8 * nrt);
    void* mem = nullptr;
    TRY(device_alloc(ctx, &mem, rows_cv.off));
    char* M = static_cast<char*>(mem);
    const VoxRows R{reinterpret_cast<uint32_t*>(M + o_tri), reinterpret_cast<uint32_t*>(M + o_u),
                    reinterpret_cast<uint32_t*>(M + o_w0), reinterpret_cast<uint32_t*>(M + o_cols),
                    reinterpret_cast<uint64_t*>(M + o_coff)};
    uint32_t* tcounts = reinterpret_cast<uint32_t*>(M + o_tcnt);
    uint64_t* toffs = reinterpret_cast<uint64_t*>(M + o_toff);
    uint64_t* trow = reinterpret_cast<uint64_t*>(M + o_trow);
    launch_row_setup(st, setup, roff, cnt, cnt + 1, n_rows, R, reinterpret_cast<uint32_t*>(M + o_rtc),
                     reinterpret_cast<uint64_t*>(M + o_rto), cnt + 2, reinterpret_cast<unsigned long long*>(cnt + 5));
    uint64_t h2[6] = {};
    if (hipMemcpyAsync(h2, cnt, sizeof(h2), hipMemcpyDeviceToHost, st) != hipSuccess) {
        device_release(ctx, mem);
        return set_error(SVDAG_ERR_HIP, "column read-back failed");
    }
    int rc = finish(ctx, "voxelize (rows)");
    if (rc == 0 && h2[5]) rc = set_error(SVDAG_ERR_CAPACITY, "a block of %llu rows spans 2^32 or more voxel columns",
                                         (unsigned long long)kVoxTileItems);
    if (rc == 0 && h2[2] > rect_columns) rc = set_error(SVDAG_ERR_INTERNAL, "row ranges exceed the rectangles");
    const uint64_t columns = h2[2];
    if (rc != 0 || columns == 0) {
        device_release(ctx, mem);
        return rc;
    }
    // One evaluation of the columns with a staging area of 16 bytes per
    // column; past the context's limit (or if it cannot be allocated), two
    // (count, then write in place).
    const uint64_t stage_bytes = tiles(columns, kVoxTileItems) * 2 * kVoxTileItems * sizeof(uint64_t);
    void* stage = nullptr;
    if (stage_bytes <= ctx->vox_stage_max && device_alloc(ctx, &stage, stage_bytes) != 0) stage = nullptr;
    if (stage)
        launch_vox_emit(st, setup, R, cnt + 1, cnt + 2, columns, depth, trow, tcounts, static_cast<uint64_t*>(stage));
    else
        launch_vox_count(st, setup, R, cnt + 1, cnt + 2, columns, depth, trow, tcounts);
    launch_scan_counts(st, tcounts, toffs, cnt + 2, kVoxTileItems, cnt + 3);
    if (hipMemcpyAsync(h, cnt, sizeof(h), hipMemcpyDeviceToHost, st) != hipSuccess)
        rc = set_error(SVDAG_ERR_HIP, "count read-back failed");
    if (rc == 0) rc = finish(ctx, "voxelize (count)");
    const uint64_t keys = h[3];
    void* out = nullptr;
    if (rc == 0 && keys) rc = device_alloc(ctx, &out, 8 * keys);
    if (rc == 0 && keys) {
        if (stage)
            launch_vox_compact(st, static_cast<const uint64_t*>(stage), tcounts, toffs, cnt + 2, columns,
                               static_cast<uint64_t*>(out));
        else
            launch_vox_write(st, setup, R, trow, cnt + 1, cnt + 2, columns, depth, toffs, static_cast<uint64_t*>(out));
        rc = hipGetLastError() == hipSuccess ? finish(ctx, "voxelize (write)")
                                             : set_error(SVDAG_ERR_HIP, "voxelizer launch failed");
    }
    device_release(ctx, stage);
    device_release(ctx, mem);
    if (rc != 0) {
        device_release(ctx, out);
        return rc;
    }
    *d_keys = static_cast<uint64_t*>(out);
    *n_keys = keys;
    return 0;
}

int svdag_free_keys(svdag_context* ctx, uint64_t* d_keys)
{
    TRY(check_ctx(ctx));
    device_release(ctx, d_keys);
    return 0;
}

int svdag_sort_keys(svdag_context* ctx, const uint64_t* d_keys, uint64_t n, int32_t bits, uint64_t* d_out)
{
    return sort_keys_entry<uint64_t>(ctx, d_keys, n, bits, d_out);
}

int svdag_sort_keys_u32(svdag_context* ctx, const uint32_t* d_keys, uint64_t n, int32_t bits, uint32_t* d_out)
{
    return sort_keys_entry<uint32_t>(ctx, d_keys, n, bits, d_out);
}

int svdag_query_voxels(svdag_context* ctx, const uint32_t* d_words, uint64_t n_words, int32_t depth,
                       const uint64_t* d_codes, uint64_t n, uint8_t* d_out)
{
    TRY(check_ctx(ctx));
    if (depth < 3 || depth > 21) return set_error(SVDAG_ERR_ARGUMENT, "depth %d outside [3, 21]", depth);
    if (n && (!d_codes || !d_out)) return set_error(SVDAG_ERR_ARGUMENT, "null codes or output");
    if (n_words && !d_words) return set_error(SVDAG_ERR_ARGUMENT, "null words");
    if (n == 0) return 0;
    launch_query_voxels(ctx->stream, d_words, n_words, depth, d_codes, n, d_out);
    HIP_TRY(hipGetLastError());
    return finish(ctx, "query_voxels");
}

int svdag_trace_params_from_view(const double* position, const double* right, const double* up,
                                 const double* direction, double fov_degrees, uint32_t width, uint32_t height,
                                 const double* aabb_min, const double* aabb_max, int32_t depth, uint32_t debug_level,
                                 svdag_trace_params* out)
{
#pragma clang fp contract(off)
    if (!position || !right || !up || !direction || !aabb_min || !aabb_max || !out)
        return set_error(SVDAG_ERR_ARGUMENT, "null argument");
    if (width == 0 || height == 0) return set_error(SVDAG_ERR_ARGUMENT, "empty image %ux%u", width, height);
    if (depth < 3 || depth > 21) return set_error(SVDAG_ERR_ARGUMENT, "depth %d outside [3, 21]", depth);
    // src/DAGTracer.h:43-80, operation for operation.
    const double fov = fov_degrees / 2.0 * (3.14159265358979323846 / 180.);
    const double aspect = double(width) / double(height);
    const double s = std::sin(fov), c = std::cos(fov);
    svdag_trace_params p{};
    p.debug_level = debug_level;
    p.width = width;
    p.height = height;
    for (int a = 0; a < 3; ++a) {
        const double x = right[a] * s * aspect, y = up[a] * s, z = direction[a] * c;
        const double bottom_left = position[a] + z - y - x;
        const double bottom_right = position[a] + z - y + x;
        const double top_left = position[a] + z + y - x;
        const double translation = -aabb_min[a];
        const double scale = double(1 << depth) / (aabb_max[a] - aabb_min[a]);
        const double f_bl = (bottom_left + translation) * scale;
        p.camera_position[a] = (position[a] + translation) * scale;
        p.ray_min[a] = f_bl;
        p.ray_ddx[a] = ((bottom_right + translation) * scale - f_bl) * (1.0 / width);
        p.ray_ddy[a] = ((top_left + translation) * scale - f_bl) * (1.0 / height);
    }
    *out = p;
    return 0;
}

int svdag_trace_paths(svdag_context* ctx, const uint32_t* d_words, uint64_t n_words, int32_t depth,
                      const svdag_trace_params* params, uint32_t* d_paths, uint32_t* d_colors)
{
    TRY(check_ctx(ctx));
    if (!params) return set_error(SVDAG_ERR_ARGUMENT, "null trace parameters");
    if (depth < 3 || depth > 21) return set_error(SVDAG_ERR_ARGUMENT, "depth %d outside [3, 21]", depth);
    if (n_words && !d_words) return set_error(SVDAG_ERR_ARGUMENT, "null words");
    if (params->width > 65535u * 16u || params->height > 65535u * 16u)
        return set_error(SVDAG_ERR_ARGUMENT, "image %ux%u too large", params->width, params->height);
    if (!d_paths && !d_colors) return 0;
    if (params->width == 0 || params->height == 0) return 0;
    launch_trace_paths(ctx->stream, d_words, n_words, depth, *params, d_paths, d_colors);
    HIP_TRY(hipGetLastError());
    return finish(ctx, "trace_paths");
}

int svdag_profile_enable(svdag_context* ctx, uint64_t kernel_mask)
{
    TRY(check_ctx(ctx));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->recorder.resolve(true);
    ctx->recorder.mask = kernel_mask & ((1ull << kNumKernelIds) - 1);
    active_recorder() = ctx->recorder.mask ? &ctx->recorder : nullptr;
    return 0;
}

int svdag_profile_reset(svdag_context* ctx)
{
    if (!ctx) return set_error(SVDAG_ERR_ARGUMENT, "null context");
    ctx->recorder.reset();
    return 0;
}

int svdag_profile_read(svdag_context* ctx, int32_t kernel, double* total_ms, uint64_t* launches)
{
    if (!ctx || !total_ms || !launches) return set_error(SVDAG_ERR_ARGUMENT, "null argument");
    if (kernel < 0 || kernel >= kNumKernelIds) return set_error(SVDAG_ERR_ARGUMENT, "kernel %d out of range", kernel);
    *total_ms = ctx->recorder.total_ms[kernel];
    *launches = ctx->recorder.launches[kernel];
    return 0;
}

const char* svdag_kernel_name(int32_t kernel)
{
    return kernel >= 0 && kernel < kNumKernelIds ? kKernelNames[kernel] : nullptr;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batched build captured as a HIP graph.  The build is enqueued once on the
// graph's own stream under stream capture, sized for max_keys keys: the chunk
// offsets are read on the device (checked there, and the segmented sort's
// tiles planned from them), every kernel takes its real counts from device
// memory, the leaf table is sized from its bound (no estimate to retry), and no
// stream is forked.  A replay therefore needs no host work: the caller writes
// new keys and offsets into the same device buffers and launches the graph.
// ---------------------------------------------------------------------------
struct svdag_batched_graph {
    svdag_context* ctx = nullptr;  // the graph's own context, on its own stream
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    char* arena = nullptr;
    char* storage = nullptr;
    BuildResult R;
    bool launched = false;
};

extern "C" {

int svdag_batched_graph_destroy(svdag_batched_graph* g)
{
    if (!g) return 0;
    if (g->ctx) (void)hipSetDevice(g->ctx->device);
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    if (g->arena) (void)hipFree(g->arena);
    if (g->storage) (void)hipFree(g->storage);
    if (g->ctx) (void)svdag_context_destroy(g->ctx);
    if (g->stream) (void)hipStreamDestroy(g->stream);
    delete g;
    return 0;
}

int svdag_batched_graph_create(svdag_context* ctx, const uint64_t* d_keys, const uint64_t* d_offsets,
                               uint64_t num_chunks, int32_t chunk_depth, uint64_t max_keys,
                               svdag_batched_graph** out)
{
    TRY(check_ctx(ctx));
    if (!out || !d_keys || !d_offsets) return set_error(SVDAG_ERR_ARGUMENT, "null argument");
    *out = nullptr;
    int extra = 0;
    for (uint64_t c = num_chunks; c > 1; c /= 8) {
        if (c % 8) return set_error(SVDAG_ERR_ARGUMENT, "chunk count %llu is not a power of 8", (unsigned long long)num_chunks);
        ++extra;
    }
    if (num_chunks == 0) return set_error(SVDAG_ERR_ARGUMENT, "no chunks");
    const int depth = chunk_depth + extra;
    if (chunk_depth < 3 || depth > 21) return set_error(SVDAG_ERR_ARGUMENT, "bad chunk depth %d", chunk_depth);
    if (max_keys == 0 || max_keys >= (1ull << 30))
        return set_error(SVDAG_ERR_ARGUMENT, "max_keys must be in [1, 2^30)");
    auto* g = new svdag_batched_graph();
    auto fail = [&](int code) {
        (void)svdag_batched_graph_destroy(g);
        return code;
    };
    if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess) {
        g->stream = nullptr;
        return fail(set_error(SVDAG_ERR_HIP, "cannot create the graph's stream"));
    }
    if (const int rc = svdag_context_create(&g->ctx, g->stream)) return fail(rc);
    g->ctx->load_shift = ctx->load_shift;
    g->ctx->radix_bits = ctx->radix_bits;
    g->ctx->sort_ipt = ctx->sort_ipt;
    g->ctx->seg_first_bits = ctx->seg_first_bits;
    g->ctx->check_collisions = ctx->check_collisions;
    g->ctx->wait_seconds = ctx->wait_seconds;
    const BatchInfo batch{nullptr, num_chunks, chunk_depth, d_offsets};
    const BuildLayout L = plan_build(max_keys, depth, 8, g->ctx->load_shift, true, g->ctx->radix_bits,
                                     g->ctx->sort_ipt);
    const uint64_t arena_bytes = L.total + plan_batch(L, max_keys, &batch).total;
    const uint64_t storage_bytes = result_bytes(max_keys, depth);
    if (ctx->budget && arena_bytes + storage_bytes > ctx->budget)
        return fail(set_error(SVDAG_ERR_MEMORY, "graph needs %llu bytes, the budget is %llu",
                              (unsigned long long)(arena_bytes + storage_bytes), (unsigned long long)ctx->budget));
    if (hipMalloc(reinterpret_cast<void**>(&g->arena), arena_bytes) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&g->storage), storage_bytes) != hipSuccess) {
        (void)hipGetLastError();
        return fail(set_error(SVDAG_ERR_MEMORY, "cannot allocate %llu bytes for the graph",
                              (unsigned long long)(arena_bytes + storage_bytes)));
    }
    FixedMemory fm{g->arena, arena_bytes, g->storage, storage_bytes};
    fm.captured = &g->R;
    active_recorder() = nullptr;  // timing events are not part of the graph
    if (hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
        return fail(set_error(SVDAG_ERR_HIP, "stream capture did not start"));
    svdag_dag unused;
    const int rc = build<uint64_t>(g->ctx, d_keys, max_keys, depth, &unused, &batch, false, &fm);
    const hipError_t ec = hipStreamEndCapture(g->stream, &g->graph);
    if (rc) return fail(rc);
    if (ec != hipSuccess || !g->graph) return fail(set_error(SVDAG_ERR_HIP, "stream capture failed: %s", hipGetErrorString(ec)));
    if (hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) != hipSuccess)
        return fail(set_error(SVDAG_ERR_HIP, "graph instantiation failed"));
    *out = g;
    return 0;
}

int svdag_batched_graph_launch(svdag_batched_graph* g)
{
    if (!g) return set_error(SVDAG_ERR_ARGUMENT, "null graph");
    TRY(check_ctx(g->ctx));
    HIP_TRY(hipGraphLaunch(g->exec, g->stream));
    g->launched = true;
    return 0;
}

int svdag_batched_graph_result(svdag_batched_graph* g, svdag_dag* out)
{
    if (!g || !out) return set_error(SVDAG_ERR_ARGUMENT, "null argument");
    TRY(check_ctx(g->ctx));
    if (!g->launched) return set_error(SVDAG_ERR_ARGUMENT, "the graph has not been launched");
    TRY(finish(g->ctx, "batched graph"));
    const int rc = complete_build(g->R, *g->ctx->h_counters, false, out);
    if (rc) return rc;
    out->storage = nullptr;  // the graph's storage, valid until its next launch
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Resident DAG (served mode).  The union of every chunk DAG inserted so far,
// kept on the device between insertions: per level one persistent ordered
// dedup table (hash -> slot) with the nodes' records stored at their slots
// (children as slots of the level below), a hashed leaf table, and per chunk
// slot its root's slot.  A node's identity is its hash, so inserting a round
// (a batched build of some chunks) adds exactly the nodes the union gains;
// the reference's MergeDAGs would give the same levels
// (src/DAGCompression.cu:1109-1320, 1447-1548).  svdag_resident_dag orders the
// levels by hash (the level extract), ranks the children, and builds the top
// levels over the filled chunk slots.
// ---------------------------------------------------------------------------
struct svdag_resident {
    svdag_context* ctx = nullptr;
    int chunk_depth = 0, top = 0, depth = 0, rlevels = 0;  // resident levels: global top .. depth - 3
    char* mem = nullptr;
    uint64_t mem_bytes = 0;
    uint64_t* tkeys[SVDAG_MAX_LEVELS] = {};
    uint32_t* trep[SVDAG_MAX_LEVELS] = {};
    uint8_t* rmasks[SVDAG_MAX_LEVELS] = {};
    uint32_t* rchildren[SVDAG_MAX_LEVELS] = {};
    uint32_t lg[SVDAG_MAX_LEVELS] = {};
    uint64_t slots[SVDAG_MAX_LEVELS] = {}, cap[SVDAG_MAX_LEVELS] = {};
    uint64_t* leaf_keys = nullptr;
    uint32_t leaf_lg = 0;
    uint64_t leaf_slots = 0, leaf_cap = 0;
    uint32_t* root_slot = nullptr;         // [8^top]
    uint64_t* root_hash = nullptr;         // [8^top]
    // Device state words: [0, rlevels] placed keys per resident level (leaves
    // last), then the error bits; then the tables' log2 sizes as u32.
    unsigned long long* state = nullptr;
    uint32_t* d_lg = nullptr;              // [rlevels + 1], leaves last
    uint64_t* h_state = nullptr;           // pinned: state words, then 8^top root hashes
    uint64_t occupied[SVDAG_MAX_LEVELS + 1] = {};
    uint64_t chunks = 0;
    bool broken = false;
};

namespace {

int resident_check(svdag_resident* r)
{
    if (!r) return set_error(SVDAG_ERR_ARGUMENT, "null resident DAG");
    TRY(check_ctx(r->ctx));
    if (r->broken) return set_error(SVDAG_ERR_INTERNAL, "resident DAG is unusable after a failed insertion");
    return 0;
}

}  // namespace

extern "C" {

int svdag_resident_create(svdag_context* ctx, int32_t chunk_depth, int32_t slot_levels, uint64_t max_nodes,
                          svdag_resident** out)
{
    TRY(check_ctx(ctx));
    if (!out) return set_error(SVDAG_ERR_ARGUMENT, "null output");
    *out = nullptr;
    if (slot_levels < 1 || slot_levels > kMaxResidentTop)
        return set_error(SVDAG_ERR_ARGUMENT, "slot levels %d outside [1, %d]", slot_levels, kMaxResidentTop);
    if (chunk_depth < 3 || chunk_depth + slot_levels > 21)
        return set_error(SVDAG_ERR_ARGUMENT, "chunk depth %d with %d slot levels out of range", chunk_depth, slot_levels);
    if (max_nodes == 0 || max_nodes >= (1ull << 31)) return set_error(SVDAG_ERR_ARGUMENT, "max_nodes must be in [1, 2^31)");
    auto* r = new svdag_resident();
    r->ctx = ctx;
    r->chunk_depth = chunk_depth;
    r->top = slot_levels;
    r->depth = chunk_depth + slot_levels;
    r->rlevels = chunk_depth - 2;
    const uint32_t shift = ctx->load_shift;
    const uint64_t nslots = 1ull << (3 * r->top);
    Carver cv;
    uint64_t o_keys[SVDAG_MAX_LEVELS], o_rep[SVDAG_MAX_LEVELS], o_masks[SVDAG_MAX_LEVELS], o_children[SVDAG_MAX_LEVELS];
    for (int i = 0; i < r->rlevels; ++i) {
        r->cap[i] = pow8_capped(r->top + i, max_nodes);
        r->lg[i] = table_lg(r->cap[i], shift);
        r->slots[i] = (1ull << r->lg[i]) + kTableOverflow;
        o_keys[i] = cv.take(8 * r->slots[i]);
        o_rep[i] = cv.take(4 * r->slots[i]);
        o_masks[i] = cv.take(r->slots[i]);
        o_children[i] = cv.take(32 * r->slots[i]);
    }
    r->leaf_cap = pow8_capped(r->depth - 2, max_nodes);
    r->leaf_lg = table_lg(r->leaf_cap, shift);
    r->leaf_slots = 1ull << r->leaf_lg;
    const uint64_t o_leaf = cv.take(8 * r->leaf_slots);
    const uint64_t o_root_slot = cv.take(4 * nslots), o_root_hash = cv.take(8 * nslots);
    const uint64_t o_state = cv.take(8 * (SVDAG_MAX_LEVELS + 2));
    const uint64_t o_lg = cv.take(4 * (SVDAG_MAX_LEVELS + 1));
    r->mem_bytes = cv.off;
    auto fail = [&](int code) {
        if (r->h_state) (void)hipHostFree(r->h_state);
        if (r->mem) (void)hipFree(r->mem);
        delete r;
        return code;
    };
    if (ctx->budget && ctx->arena_cap + ctx->pool_bytes + r->mem_bytes > ctx->budget)
        return fail(set_error(SVDAG_ERR_MEMORY, "resident tables of %llu bytes exceed the memory budget",
                              (unsigned long long)r->mem_bytes));
    if (hipMalloc(reinterpret_cast<void**>(&r->mem), r->mem_bytes) != hipSuccess) {
        (void)hipGetLastError();
        r->mem = nullptr;
        return fail(set_error(SVDAG_ERR_MEMORY, "cannot allocate %llu bytes of resident tables",
                              (unsigned long long)r->mem_bytes));
    }
    if (hipHostMalloc(reinterpret_cast<void**>(&r->h_state), 8 * (SVDAG_MAX_LEVELS + 2 + nslots),
                      hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        r->h_state = nullptr;
        return fail(set_error(SVDAG_ERR_HIP, "cannot allocate pinned staging"));
    }
    char* M = r->mem;
    for (int i = 0; i < r->rlevels; ++i) {
        r->tkeys[i] = reinterpret_cast<uint64_t*>(M + o_keys[i]);
        r->trep[i] = reinterpret_cast<uint32_t*>(M + o_rep[i]);
        r->rmasks[i] = reinterpret_cast<uint8_t*>(M + o_masks[i]);
        r->rchildren[i] = reinterpret_cast<uint32_t*>(M + o_children[i]);
    }
    r->leaf_keys = reinterpret_cast<uint64_t*>(M + o_leaf);
    r->root_slot = reinterpret_cast<uint32_t*>(M + o_root_slot);
    r->root_hash = reinterpret_cast<uint64_t*>(M + o_root_hash);
    r->state = reinterpret_cast<unsigned long long*>(M + o_state);
    r->d_lg = reinterpret_cast<uint32_t*>(M + o_lg);
    // Empty tables, no chunk slot filled, zero counts; the log2 sizes as
    // device words (the extract reads them there).
    hipStream_t st = ctx->stream;
    for (int i = 0; i < r->rlevels; ++i) launch_fill_u64(st, r->tkeys[i], r->slots[i], 0);
    launch_fill_u64(st, r->leaf_keys, r->leaf_slots, 0);
    launch_fill_u64(st, reinterpret_cast<uint64_t*>(r->root_slot), (nslots + 1) / 2, ~0ull);
    launch_fill_u64(st, reinterpret_cast<uint64_t*>(r->state), SVDAG_MAX_LEVELS + 2, 0);
    std::vector<uint32_t> lgs(SVDAG_MAX_LEVELS + 1, 0);
    for (int i = 0; i < r->rlevels; ++i) lgs[i] = r->lg[i];
    lgs[r->rlevels] = r->leaf_lg;
    if (hipMemcpyAsync(r->d_lg, lgs.data(), 4 * lgs.size(), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return fail(set_error(SVDAG_ERR_HIP, "resident table set-up failed"));
    if (hipGetLastError() != hipSuccess) return fail(set_error(SVDAG_ERR_HIP, "resident table set-up failed"));
    *out = r;
    return 0;
}

int svdag_resident_destroy(svdag_resident* r)
{
    if (!r) return 0;
    if (r->ctx) {
        (void)hipSetDevice(r->ctx->device);
        (void)hipStreamSynchronize(r->ctx->stream);
    }
    if (r->mem) (void)hipFree(r->mem);
    if (r->h_state) (void)hipHostFree(r->h_state);
    delete r;
    return 0;
}

int svdag_resident_insert(svdag_resident* r, const svdag_dag* round, uint64_t* root_hashes)
{
    TRY(resident_check(r));
    svdag_context* ctx = r->ctx;
    if (!round) return set_error(SVDAG_ERR_ARGUMENT, "null round");
    if (round->num_levels != r->depth - 2)
        return set_error(SVDAG_ERR_ARGUMENT, "round has %d levels, the resident DAG %d", round->num_levels,
                         r->depth - 2);
    const uint64_t nslots = 1ull << (3 * r->top);
    if (root_hashes) std::memset(root_hashes, 0, 8 * nslots);
    if (round->num_leaves == 0) return 0;
    if (round->location != SVDAG_DEVICE) return set_error(SVDAG_ERR_ARGUMENT, "round must be device-resident");
    if (round->levels[0].num_nodes != 1) return set_error(SVDAG_ERR_ARGUMENT, "round has no single root");
    // Admission: the worst case (every node of the round new) must fit each table.
    for (int i = 0; i < r->rlevels; ++i)
        if (r->occupied[i] + round->levels[r->top + i].num_nodes > r->cap[i])
            return set_error(SVDAG_ERR_MEMORY, "resident level %d would exceed its capacity of %llu nodes", r->top + i,
                             (unsigned long long)r->cap[i]);
    if (r->occupied[r->rlevels] + round->num_leaves > r->leaf_cap)
        return set_error(SVDAG_ERR_MEMORY, "resident leaves would exceed their capacity of %llu",
                         (unsigned long long)r->leaf_cap);
    uint64_t widest = round->num_leaves;
    for (int i = 0; i < r->rlevels; ++i) widest = std::max(widest, round->levels[r->top + i].num_nodes);
    Carver cv;
    const uint64_t o_map0 = cv.take(4 * widest), o_map1 = cv.take(4 * widest);
    TRY(ensure_arena(ctx, cv.off));
    uint32_t* map[2] = {reinterpret_cast<uint32_t*>(ctx->arena + o_map0), reinterpret_cast<uint32_t*>(ctx->arena + o_map1)};
    hipStream_t st = ctx->stream;
    unsigned long long* err = r->state + SVDAG_MAX_LEVELS + 1;
    // Leaves, then levels bottom-up: each round node's resident slot is the
    // child map of the level above.
    launch_resident_leaves(st, round->leaves, round->num_leaves, r->leaf_keys, r->leaf_lg, map[0],
                           r->state + r->rlevels, err);
    int cur = 0;
    for (int i = r->rlevels - 1; i >= 0; --i) {
        const svdag_level& lv = round->levels[r->top + i];
        ResidentLevelArgs a{};
        a.hashes = lv.hashes;
        a.masks = lv.masks;
        a.children = lv.children;
        a.count = lv.num_nodes;
        a.child_slot_of = map[cur];
        a.table_keys = r->tkeys[i];
        a.table_rep = r->trep[i];
        a.lg = r->lg[i];
        a.rec_masks = r->rmasks[i];
        a.rec_children = r->rchildren[i];
        a.slot_of = map[cur ^ 1];
        a.occupied = r->state + i;
        a.err = err;
        launch_resident_level(st, a);
        cur ^= 1;
    }
    ResidentRootArgs ra{};
    ra.top_levels = r->top;
    ra.has_root = true;
    for (int l = 0; l < r->top; ++l) {
        ra.masks[l] = round->levels[l].masks;
        ra.children[l] = round->levels[l].children;
    }
    ra.root_slot_of = map[cur];
    ra.root_table_keys = r->tkeys[0];
    ra.root_slot = r->root_slot;
    ra.root_hash = r->root_hash;
    ra.err = err;
    launch_resident_roots(st, ra);
    if (hipGetLastError() != hipSuccess) {
        r->broken = true;
        return set_error(SVDAG_ERR_HIP, "resident insert launch failed");
    }
    if (hipMemcpyAsync(r->h_state, r->state, 8 * (SVDAG_MAX_LEVELS + 2), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(r->h_state + SVDAG_MAX_LEVELS + 2, r->root_hash, 8 * nslots, hipMemcpyDeviceToHost, st) !=
            hipSuccess) {
        r->broken = true;
        return set_error(SVDAG_ERR_HIP, "resident state read-back failed");
    }
    mark(ctx, "resident insert");
    if (const int rc = finish(ctx, "resident_insert")) {
        r->broken = true;
        return rc;
    }
    const uint64_t e = r->h_state[SVDAG_MAX_LEVELS + 1];
    if (e) {
        r->broken = true;
        if (e & kErrSlotFilled) return set_error(SVDAG_ERR_ARGUMENT, "a chunk of the round is at a slot already filled");
        return set_error(SVDAG_ERR_MEMORY, "resident table probe overflow");
    }
    for (int i = 0; i <= r->rlevels; ++i) r->occupied[i] = r->h_state[i];
    const uint64_t* hashes = r->h_state + SVDAG_MAX_LEVELS + 2;
    for (uint64_t s = 0; s < nslots; ++s) r->chunks += hashes[s] != 0;
    if (root_hashes) std::memcpy(root_hashes, hashes, 8 * nslots);
    return 0;
}

int svdag_resident_counts(svdag_resident* r, uint64_t* nodes, uint64_t* chunks)
{
    if (!r) return set_error(SVDAG_ERR_ARGUMENT, "null resident DAG");
    if (nodes)
        for (int i = 0; i <= r->rlevels; ++i) nodes[i] = r->occupied[i];
    if (chunks) *chunks = r->chunks;
    return 0;
}

int svdag_resident_dag(svdag_resident* r, svdag_dag* out)
{
    TRY(resident_check(r));
    if (!out) return set_error(SVDAG_ERR_ARGUMENT, "null output");
    svdag_context* ctx = r->ctx;
    const int levels = r->depth - 2;
    empty_dag(out, levels, SVDAG_DEVICE);
    if (r->chunks == 0) return 0;
    hipStream_t st = ctx->stream;
    const uint64_t nslots = 1ull << (3 * r->top);
    const uint64_t L = r->occupied[r->rlevels];
    const SortPlan leaf_plan = make_sort_plan(0, 64);
    uint64_t max_slots = r->leaf_slots;
    for (int i = 0; i < r->rlevels; ++i) max_slots = std::max(max_slots, r->slots[i]);
    const uint64_t stage_tiles = std::max({tiles(max_slots, kExtractTileItems), tiles(r->leaf_slots, kListTileItems),
                                           uint64_t(1)});
    Carver cv;
    const uint64_t o_zero = cv.off;
    const uint64_t o_counters = cv.take(sizeof(BuildCounters));
    const uint64_t o_sort_zero = cv.take(4 * sort_zero_words(leaf_plan));
    const uint64_t zero_end = cv.off;
    const uint64_t o_sort_counts = cv.take(4 * sort_count_words(L));
    const uint64_t o_tile_counts = cv.take(4 * stage_tiles), o_tile_offsets = cv.take(8 * stage_tiles);
    uint64_t o_lk[2], o_ls[2];
    for (int i = 0; i < 2; ++i) {
        o_lk[i] = cv.take(8 * L);
        o_ls[i] = cv.take(4 * L);
    }
    const uint64_t o_sorted = cv.take(4 * L), o_leaf_rank = cv.take(4 * r->leaf_slots);
    const uint64_t o_rank[2] = {cv.take(4 * max_slots), cv.take(4 * max_slots)};
    const uint64_t o_codes = cv.take(8 * nslots), o_rslots = cv.take(4 * nslots), o_rcount = cv.take(8);
    uint64_t o_tcodes[2], o_tranks[2];
    for (int i = 0; i < 2; ++i) {
        o_tcodes[i] = cv.take(8 * kTailMaxParents);
        o_tranks[i] = cv.take(4 * kTailMaxParents);
    }
    const uint64_t o_tmasks = cv.take(kTailMaxParents), o_tchildren = cv.take(32 * kTailMaxParents);
    const uint64_t o_thash = cv.take(64 * kTailMaxParents);
    TRY(ensure_arena(ctx, cv.off));
    char* A = ctx->arena;
    BuildCounters* c = reinterpret_cast<BuildCounters*>(A + o_counters);
    uint32_t* tile_counts = reinterpret_cast<uint32_t*>(A + o_tile_counts);
    uint64_t* tile_offsets = reinterpret_cast<uint64_t*>(A + o_tile_offsets);

    uint64_t caps[SVDAG_MAX_LEVELS];
    for (int l = 0; l < levels; ++l) caps[l] = l < r->top ? (1ull << (3 * l)) : r->occupied[l - r->top];
    const DagStorage S = plan_storage(levels, caps, L);
    void* storage = nullptr;
    TRY(device_alloc(ctx, &storage, S.total));
    char* O = static_cast<char*>(storage);
    auto fail = [&](int code) {
        device_release(ctx, storage);
        return code;
    };
    launch_fill_u64(st, reinterpret_cast<uint64_t*>(A + o_zero), (zero_end - o_zero) / 8, 0);
    // Leaves in ascending mask order, and each leaf slot's rank.
    uint64_t* lk0 = reinterpret_cast<uint64_t*>(A + o_lk[0]);
    uint64_t* lk1 = reinterpret_cast<uint64_t*>(A + o_lk[1]);
    uint32_t* ls0 = reinterpret_cast<uint32_t*>(A + o_ls[0]);
    uint32_t* ls1 = reinterpret_cast<uint32_t*>(A + o_ls[1]);
    uint32_t* sorted = reinterpret_cast<uint32_t*>(A + o_sorted);
    uint32_t* leaf_rank = reinterpret_cast<uint32_t*>(A + o_leaf_rank);
    uint64_t* leaves_out = reinterpret_cast<uint64_t*>(O + S.leaves);
    launch_slot_list(st, r->leaf_keys, r->d_lg + r->rlevels, r->leaf_slots, tile_counts, tile_offsets,
                     &c->table_slots, lk0, ls0, &c->unique_leaves);
    if (launch_sort<uint64_t>(st, leaf_plan, lk0, ls0, lk1, lk0, ls1, ls0, leaves_out, sorted, &c->unique_leaves, L,
                              reinterpret_cast<uint32_t*>(A + o_sort_zero),
                              reinterpret_cast<uint32_t*>(A + o_sort_counts), nullptr) < 0)
        return fail(set_error(SVDAG_ERR_INTERNAL, "leaf sort plan failed"));
    launch_rank_of_slot(st, sorted, &c->unique_leaves, leaf_rank, L);
    // Resident levels bottom-up, each in ascending hash order with its
    // children's slots turned into ranks of the level below.
    const uint32_t* below = leaf_rank;
    int cur = 0;
    for (int i = r->rlevels - 1; i >= 0; --i, cur ^= 1) {
        const int l = r->top + i;
        LevelExtractArgs e{};
        e.table_keys = r->tkeys[i];
        e.table_rep = r->trep[i];
        e.table_log2 = r->d_lg + i;
        e.rank_of_slot = reinterpret_cast<uint32_t*>(A + o_rank[cur]);
        e.parent_masks = r->rmasks[i];
        e.parent_children = r->rchildren[i];
        e.child_rank_of_slot = below;
        e.out_hashes = reinterpret_cast<uint64_t*>(O + S.hashes[l]);
        e.out_masks = reinterpret_cast<uint8_t*>(O + S.masks[l]);
        e.out_children = reinterpret_cast<uint32_t*>(O + S.children[l]);
        e.tile_counts = tile_counts;
        e.offsets = tile_offsets;
        e.slots_scratch = &c->table_slots;
        e.q_count = &c->uniq[l];
        e.w_count = &c->words[l];
        launch_level_extract(st, e, r->slots[i]);
        below = e.rank_of_slot;
    }
    // Top levels over the filled chunk slots (their codes are the slot indices).
    uint64_t* codes = reinterpret_cast<uint64_t*>(A + o_codes);
    uint32_t* rslots = reinterpret_cast<uint32_t*>(A + o_rslots);
    uint64_t* rcount = reinterpret_cast<uint64_t*>(A + o_rcount);
    launch_resident_root_list(st, r->root_slot, uint32_t(nslots), codes, rslots, rcount);
    TailArgs t{};
    t.top = r->top - 1;
    t.child_codes = codes;
    t.child_slots = rslots;
    t.child_rank_of_slot = below;
    t.child_hashes = reinterpret_cast<const uint64_t*>(O + S.hashes[r->top]);
    t.child_count = rcount;
    for (int i = 0; i < 2; ++i) {
        t.codes[i] = reinterpret_cast<uint64_t*>(A + o_tcodes[i]);
        t.ranks[i] = reinterpret_cast<uint32_t*>(A + o_tranks[i]);
    }
    t.masks = reinterpret_cast<uint8_t*>(A + o_tmasks);
    t.children = reinterpret_cast<uint32_t*>(A + o_tchildren);
    t.child_hash = reinterpret_cast<uint64_t*>(A + o_thash);
    for (int l = 0; l < r->top; ++l) {
        t.out_hashes[l] = reinterpret_cast<uint64_t*>(O + S.hashes[l]);
        t.out_masks[l] = reinterpret_cast<uint8_t*>(O + S.masks[l]);
        t.out_children[l] = reinterpret_cast<uint32_t*>(O + S.children[l]);
    }
    t.counters = c;
    t.check_collisions = 0;
    launch_level_tail(st, t);
    if (hipGetLastError() != hipSuccess) return fail(set_error(SVDAG_ERR_HIP, "resident extract launch failed"));
    if (hipMemcpyAsync(ctx->h_counters, c, sizeof(BuildCounters), hipMemcpyDeviceToHost, st) != hipSuccess)
        return fail(set_error(SVDAG_ERR_HIP, "counter read-back failed"));
    mark(ctx, "resident extract");
    if (const int rc = finish(ctx, "resident_dag")) return fail(rc);
    const BuildCounters& h = *ctx->h_counters;
    if (h.uniq[0] != 1) return fail(set_error(SVDAG_ERR_INTERNAL, "resident root level holds %llu nodes", (unsigned long long)h.uniq[0]));
    out->storage = storage;
    out->leaves = leaves_out;
    out->num_leaves = h.unique_leaves;
    for (int l = 0; l < levels; ++l) {
        out->levels[l].hashes = reinterpret_cast<uint64_t*>(O + S.hashes[l]);
        out->levels[l].children = reinterpret_cast<uint32_t*>(O + S.children[l]);
        out->levels[l].masks = reinterpret_cast<uint8_t*>(O + S.masks[l]);
        out->levels[l].num_nodes = h.uniq[l];
        out->level_words[l] = h.words[l];
    }
    return 0;
}

}  // extern "C"
