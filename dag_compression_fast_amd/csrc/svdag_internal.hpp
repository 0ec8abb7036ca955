// svdag_internal.hpp -- host-side launchers and device-memory layouts shared by
// the kernels (svdag_sort.hip, svdag_build.hip) and the C-ABI driver
// (svdag_api.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "svdag.h"

namespace svdag {

// ---- per-kernel timing ----------------------------------------------------
// Counterpart of the reference's per-operation event timer
// (src/CudaPerfRecorder.h:5-60, used at src/DAGCompression.cu:352, 384, 413...),
// which synchronises after every operation.  Here a launch is bracketed by two
// events on its stream when a recorder is active on the calling thread; the
// pairs are resolved after the call's single synchronisation.
enum KernelId : int {
    kKHistogram = 0,
    kKSortCount,
    kKSortScan,
    kKSortScatterKeys,
    kKSortScatterPairs,
    kKFill,
    kKScanCounts,
    kKLeafReduce,
    kKLeafCompact,
    kKTableClear,
    kKLeafInsert,
    kKRankOfSlot,
    kKLevelCount,
    kKLevelBuild,
    kKLevelCheck,
    kKExtractCount,
    kKLevelExtract,
    kKFinalCount,
    kKFinalPositions,
    kKFinalWrite,
    kKMergeInsert,
    kKSketch,
    kKLevelTail,
    kKSlotList,
    kKQueryVoxels,
    kKVoxSetup,
    kKVoxCount,
    kKVoxWrite,
    kKTracePaths,
    kKResidentInsert,
    kKSortScatterTop,  // top-digit pass of a top-digit-first key sort (64-bit in, 32-bit out)
    kKSegCount,        // segmented sort: per-tile counts
    kKSegScatter,      // segmented sort: scatter passes
    kKPairHistogram,   // key/value sorts (the unique-leaf order): first read
    kKPairCount,       // key/value sorts: per-tile counts of the later passes
    kNumKernelIds
};

struct KernelRecorder {
    virtual void begin(hipStream_t s, int id) = 0;
    virtual void end(hipStream_t s, int id) = 0;
    virtual ~KernelRecorder() = default;
};

// The recorder of the calling thread (null: no timing).
KernelRecorder*& active_recorder();

struct TimedLaunch {
    int id;
    hipStream_t s;
    TimedLaunch(int kernel, hipStream_t stream) : id(kernel), s(stream)
    {
        if (KernelRecorder* r = active_recorder()) r->begin(s, id);
    }
    ~TimedLaunch()
    {
        if (KernelRecorder* r = active_recorder()) r->end(s, id);
    }
};

// ---- radix sort -----------------------------------------------------------
// Digits are at most kMaxRadixBits wide; kRadix digit slots are reserved in
// every LDS and count array.
constexpr int kMaxRadixBits = 9;
constexpr int kRadix = 1 << kMaxRadixBits;
constexpr int kMaxPasses = 8;
// A tile is 256 threads x ipt keys: 16 (4096 keys; key/value sorts) or 32
// (8192 keys; keys-only sorts).
constexpr int kSortThreads = 256;
constexpr int kSortIptPairs = 8;   // u64 key/value sorts
constexpr int kSortIptSmall = 16;
constexpr int kSortIptLarge = 32;

struct SortPlan {
    int passes;
    uint32_t shift[kMaxPasses];
    uint32_t bits[kMaxPasses];
    int ipt;  // keys per thread and tile
    // Key-only sorts rank their first pass unstably unless this is set; keys
    // equal on the sorted bits then keep their input order (a stable sort).
    bool stable;
};

// Digits of width <= max_bits covering bits [lo, hi), spread evenly.
SortPlan make_sort_plan(int lo, int hi, int max_bits = kMaxRadixBits, int ipt = kSortIptSmall);

inline uint64_t sort_tiles(uint64_t n, int ipt)
{
    const uint64_t tile = uint64_t(kSortThreads) * ipt;
    return (n + tile - 1) / tile;
}

// Device scratch of one sort: the all-pass digit histogram (passes * kRadix
// u32, zero before launch_sort) and the per-tile digit counts
// (sort_count_words u32 for any tile size, any contents).
inline uint64_t sort_zero_words(const SortPlan& p) { return uint64_t(p.passes) * kRadix; }
inline uint64_t sort_count_words(uint64_t n_bound) { return uint64_t(kRadix) * sort_tiles(n_bound, kSortIptPairs); }

// LSD radix sort of keys (and optionally u32 values) on the plan's bits.  With
// values, or with plan.stable, keys equal on those bits keep their input order;
// otherwise their order is unspecified (the first pass ranks unstably).
// n is read from device memory (*n_ptr, < 2^30); n_bound is a host upper bound
// sizing the grid and the count buffer.  Passes ping-pong between (keys_a,
// keys_b); the first pass reads keys_in.  Returns which buffer holds the result
// (0 = a, 1 = b, 2 = final); if final_keys / final_vals are non-null the last
// pass writes there instead.  key_or (optional) receives the OR of all keys.
// Chunk-local input of a batched build: key i, with offsets[c] <= i <
// offsets[c + 1], becomes (c << chunk_bits) | key when the sort first reads it
// (chunk order = Morton order of the sub-volumes, src/Utils/Aabb.h:83-105).
struct ChunkPrefix {
    const uint64_t* offsets = nullptr;  // device, num_chunks + 1 entries; null: keys are global
    uint64_t num_chunks = 0;
    uint32_t chunk_bits = 0;
    unsigned long long* err = nullptr;  // receives kErrKeyRange for a key >= 2^chunk_bits
};

// Segmented key sort of a batched build (chunks back to back, in chunk order):
// tiles of at most kSegTileItems keys, none crossing a chunk, planned on the
// device from the chunk offsets (chunk_start: num_chunks + 1 device entries;
// *n_ptr keys in all, at most n_bound; *n_ptr == 0 plans nothing).  Sorts the
// chunk-local bits [lo, chunk_bits) (at most kSegMaxBits of them, in one or
// two passes of at most kMaxSegDigitBits) and writes the keys as 32-bit
// chunk-local keys (chunk_bits <= 32); a reader restores key i's chunk index
// from chunk_start (ChunkPrefix).  Returns 0 (result in keys_a), 1 (keys_b),
// or -1 when the bit span does not fit.  first_bits: the first pass's digit width (0: half the
// bits), clamped so both passes have at most kMaxSegDigitBits.
struct SegTile {
    uint64_t start;
    uint32_t count;
    uint32_t chunk;
};
constexpr uint32_t kSegTileItems = 4096;
constexpr uint32_t kSegMaxBits = 12;
constexpr uint32_t kMaxSegDigitBits = 9;
inline uint64_t seg_tile_bound(uint64_t n_bound, uint64_t num_chunks) { return n_bound / kSegTileItems + num_chunks; }
// When plain builds of 64-bit keys sort top digit first (option
// SVDAG_OPT_TOP_DIGIT_FIRST): never, from kTopFirstMinKeys keys on, always
// (where the bits fit).
constexpr int kTopFirstNever = 0, kTopFirstAuto = 1, kTopFirstAlways = 2;
constexpr uint64_t kTopFirstMinKeys = 1ull << 20;
// Scratch of the segmented sort: tiles (seg_tile_bound entries), chunk_tiles
// (num_chunks + 1), counts (tile bound << kMaxSegDigitBits), totals and starts
// (num_chunks << kMaxSegDigitBits each), the tile count.
struct SegScratch {
    SegTile* tiles;
    uint32_t* chunk_tiles;
    uint32_t* counts;
    uint32_t* totals;
    uint32_t* starts;
    uint64_t* ntiles;
};
template <typename KI>
int launch_segmented_sort(hipStream_t stream, const KI* keys_in, uint32_t* keys_a, uint32_t* keys_b,
                          const uint64_t* chunk_start, uint64_t num_chunks, const uint64_t* n_ptr, uint64_t n_bound,
                          const SegScratch& sc, uint32_t lo, uint32_t chunk_bits, unsigned long long* err,
                          uint32_t first_bits = 0);

// Key sort of a plain build of 64-bit keys on bits [lo, top_shift + top_bits),
// top digit first: a pass on the top digit writes each key's bits below
// top_shift (< 2^32) into `local`, grouped by top digit, with the groups'
// starts in bucket_start (2^top_bits + 1 entries); the segmented sort then
// sorts [lo, top_shift) inside each group, still as 32-bit bucket-local keys.  zero:
// kRadix u32, zero at launch; counts: sort_count_words(n_bound) u32.  Returns
// as launch_segmented_sort (0: keys_a, 1: keys_b), -1 when the bits do not fit
// (top_shift > 32, or more than kSegMaxBits below it).
int launch_top_digit_sort(hipStream_t stream, const uint64_t* keys_in, uint32_t* local, uint32_t* keys_a,
                          uint32_t* keys_b, int ipt, uint32_t top_shift, uint32_t top_bits, uint32_t lo,
                          const uint64_t* n_ptr, uint64_t n_bound, uint32_t* zero, uint32_t* counts,
                          unsigned long long* key_or, uint64_t* bucket_start, const SegScratch& sc,
                          unsigned long long* err);

template <typename K>
int launch_sort(hipStream_t stream, const SortPlan& plan, const K* keys_in, const uint32_t* vals_in,
                K* keys_a, K* keys_b, uint32_t* vals_a, uint32_t* vals_b, K* final_keys, uint32_t* final_vals,
                const uint64_t* n_ptr, uint64_t n_bound, uint32_t* zero, uint32_t* counts,
                unsigned long long* key_or, const ChunkPrefix& prefix = ChunkPrefix{});

// ---- build ----------------------------------------------------------------
constexpr uint32_t kTableOverflow = 4096;  // slots past 2^log2 for ordered tables

// Device counters of one build; zeroed at the start.
struct BuildCounters {
    uint64_t n;
    unsigned long long key_or;
    uint64_t leaf_positions;                  // L0
    uint64_t unique_leaves;                   // L
    uint64_t pre[SVDAG_MAX_LEVELS];           // M_l
    uint64_t uniq[SVDAG_MAX_LEVELS];          // Q_l
    uint64_t words[SVDAG_MAX_LEVELS];         // W_l
    uint64_t collisions;
    uint64_t error;                           // kErr* bits
    uint32_t table_log2[SVDAG_MAX_LEVELS + 1];  // [l] level tables, [MAX] leaf table
    uint64_t table_slots;                     // slots of the table being extracted
    uint64_t leaf_estimate;                   // table sizing count of the leaf dedup table
};

constexpr uint64_t kErrTableOverflow = 1;
constexpr uint64_t kErrKeyRange = 2;
constexpr uint64_t kErrInternal = 4;  // a bound the kernels rely on did not hold
constexpr uint64_t kErrSlotFilled = 8;  // a resident chunk slot received a second chunk
constexpr uint64_t kErrOffsets = 16;    // device chunk offsets invalid or past the key capacity

// Tile sizes of the compaction stages; a stage over at most n items needs
// tiles(n, kXxxTile) u32 counts and u64 offsets of scratch.
constexpr uint64_t kPackTileItems = 4096;     // grouped keys per leaf-reduce chunk
constexpr uint32_t kLeafGroupBits = 9;        // low leaf-code bits the key sort leaves unsorted
constexpr uint64_t kLeafRegionItems = kPackTileItems + (1u << kLeafGroupBits);  // leaves one chunk may emit
constexpr uint64_t kBuildTileItems = 1024;    // child codes per level-build tile
constexpr uint64_t kExtractTileItems = 256;   // table slots per extract tile
constexpr uint64_t kFinalTileItems = 4096;    // nodes per final-layout tile
constexpr uint64_t kListTileItems = 4096;     // table slots per slot-list tile
inline uint64_t tiles(uint64_t n, uint64_t per) { return (n + per - 1) / per; }

int device_cu_count();

// Workgroups of `threads` threads of `kernel` that the device holds at once
// (its occupancy times the CU count; cached per kernel).  A kernel whose
// workgroups each walk several tiles launches at most this many: a larger grid
// leaves a partial second round of workgroups, each with a full share of tiles.
int kernel_slots(const void* kernel, int threads);
template <typename F>
inline int kernel_slots(F kernel, int threads)
{
    return kernel_slots(reinterpret_cast<const void*>(kernel), threads);
}
// Grid for `tiles` units of work: one workgroup per unit, at most kernel_slots.
template <typename F>
inline int resident_grid(F kernel, int threads, uint64_t tiles)
{
    return int(std::max<uint64_t>(1, std::min<uint64_t>(tiles, uint64_t(kernel_slots(kernel, threads)))));
}

// p[0, count) = value, except p[set_at] = set_value when set_at < count.
void launch_fill_u64(hipStream_t s, uint64_t* p, uint64_t count, uint64_t value, uint64_t set_at = ~0ull,
                     uint64_t set_value = 0);

// *n_out = offsets[num_chunks] when the device offsets start at 0, never
// decrease and end at most at n_bound; else *n_out = 0 and kErrOffsets in *err.
void launch_batch_count(hipStream_t s, const uint64_t* offsets, uint64_t num_chunks, uint64_t n_bound, uint64_t* n_out,
                        unsigned long long* err);

// One-workgroup exclusive scan of tile counts into offsets; the number of tiles
// is ceil(items / per_tile), items = *items_dev, or items_host when items_dev is
// null; the total goes to *total_out when non-null.
void launch_scan_counts(hipStream_t s, const uint32_t* counts, uint64_t* offsets, const uint64_t* items_dev,
                        uint64_t per_tile, uint64_t* total_out, uint64_t items_host = 0);


// Leaf ReduceByKey over keys grouped by code >> (6 + group_bits): per-chunk
// leaves into the regions reg_code / reg_mask (kLeafRegionItems per chunk of
// kPackTileItems keys), then compacted into lp_code / lp_mask (lp_capacity
// entries); sets
// c->leaf_positions and adds the leaf level's parent count to c->pre[leaf_level].
// buckets.offsets non-null (32-bit keys only): the keys are local to the
// buckets of a segmented sort, key i's code is (b << buckets.chunk_bits) | key
// for the bucket b holding position i.
template <typename K>
void launch_leaf_reduce(hipStream_t s, const K* sorted, BuildCounters* c, uint32_t group_bits, uint64_t* reg_code,
                        uint64_t* reg_mask, uint64_t* lp_code, uint64_t* lp_mask, uint64_t lp_capacity,
                        uint32_t* tile_counts, uint64_t* tile_offsets, uint64_t n_bound, int leaf_level,
                        uint64_t code_mask, const ChunkPrefix& buckets = ChunkPrefix{});

// Clears an ordered (load_shift, overflow) or hashed table sized from *count_ptr
// and records its log2 size.
void launch_table_clear(hipStream_t s, uint64_t* keys, const uint64_t* count_ptr, uint32_t* log2_out,
                        uint32_t load_shift, uint32_t overflow, uint64_t slot_bound);

// Inserts every leaf mask into the hashed table and records its slot.
void launch_leaf_insert(hipStream_t s, const uint64_t* lp_mask, const uint64_t* count_ptr, uint64_t* table_keys,
                        const uint32_t* log2_ptr, uint32_t* lp_slot, uint64_t n_bound, unsigned long long* err);

// The occupied slots of a hashed table (2^*log2_ptr slots, at most slot_bound)
// as a list (key, slot index); their number goes to *list_count.  tile_counts
// and tile_offsets need tiles(slot_bound, 4096) entries; *slots_scratch
// receives the slot count.
void launch_slot_list(hipStream_t s, const uint64_t* table_keys, const uint32_t* log2_ptr, uint64_t slot_bound,
                      uint32_t* tile_counts, uint64_t* tile_offsets, uint64_t* slots_scratch, uint64_t* list_keys,
                      uint32_t* list_slots, uint64_t* list_count);

// Distinct-count sketch (HyperLogLog, kSketchRegs registers, zero before the
// first launch) of values[0, *count_ptr), and the count a hashed table is sized
// from: 1.25 x the estimate, at most *bound_ptr.
constexpr int kSketchBits = 12;
constexpr int kSketchRegs = 1 << kSketchBits;
void launch_sketch(hipStream_t s, const uint64_t* values, const uint64_t* count_ptr, uint32_t* regs, uint64_t n_bound);
void launch_sketch_estimate(hipStream_t s, const uint32_t* regs, const uint64_t* bound_ptr, uint64_t* estimate);

void launch_rank_of_slot(hipStream_t s, const uint32_t* sorted_slots, const uint64_t* count_ptr,
                         uint32_t* rank_of_slot, uint64_t n_bound);

// Children come in one of two forms.  With child_hash_seq, child i's hash is
// child_hash_seq[i] and its table slot is recorded as the pointer (the level's
// extract turns slots into ranks); otherwise its rank is
// child_rank_of_slot[child_slots[i]] and its hash child_hashes[rank].
struct LevelBuildArgs {
    const uint64_t* child_codes;
    const uint32_t* child_slots;
    const uint64_t* child_hash_seq;
    const uint32_t* child_rank_of_slot;
    const uint64_t* child_hashes;
    const uint64_t* child_count;
    uint64_t* parent_hashes;  // hash of parent p at p, or null
    uint64_t* parent_codes;
    uint32_t* parent_slots;
    uint8_t* parent_masks;
    uint32_t* parent_children;
    uint64_t* table_keys;
    uint32_t* table_rep;
    const uint32_t* table_log2;
    uint32_t* tile_counts;    // tiles(child_bound, kBuildTileItems)
    const uint64_t* offsets;  // same count, written by the scan
    uint64_t* next_count;     // M of the level above (atomic adds), or null
    BuildCounters* counters;
};
// The level's table clear (launch_table_clear's arguments), done by the
// level's count kernel instead of a launch of its own.
struct TableClearArgs {
    uint64_t* keys;
    const uint64_t* count;  // the level's node count (sizes the table)
    uint32_t* log2_out;
    uint32_t load_shift;
    uint32_t overflow;
    uint64_t slot_bound;
};
void launch_level_build(hipStream_t s, const LevelBuildArgs& a, uint64_t child_bound,
                        const TableClearArgs* clear = nullptr);

// ---- small top levels -------------------------------------------------------
// Levels 0..top with top <= kTailTopLevel hold at most 8^top <= 512 nodes; one
// workgroup builds, dedups (LDS sort by hash) and writes all of them.
constexpr int kTailTopLevel = 3;
constexpr uint32_t kTailMaxParents = 1u << (3 * kTailTopLevel);
constexpr uint32_t kTailMaxChildren = 1u << (3 * (kTailTopLevel + 1));

struct TailArgs {
    int top;                              // highest level built here
    // Children of level `top`: codes (sorted), table slots and the child level's
    // slot -> rank map, hashes by rank, and their count.
    const uint64_t* child_codes;
    const uint32_t* child_slots;
    const uint32_t* child_rank_of_slot;
    const uint64_t* child_hashes;
    const uint64_t* child_count;
    // Scratch, each at least kTailMaxParents entries (two ping-pong halves for
    // codes and ranks).
    uint64_t* codes[2];
    uint32_t* ranks[2];
    uint8_t* masks;
    uint32_t* children;                   // 8 per parent
    uint64_t* child_hash;                 // 8 per parent
    // Output levels 0..top.
    uint64_t* out_hashes[kTailTopLevel + 1];
    uint8_t* out_masks[kTailTopLevel + 1];
    uint32_t* out_children[kTailTopLevel + 1];
    BuildCounters* counters;
    int check_collisions;
};
void launch_level_tail(hipStream_t s, const TailArgs& a);

struct LevelExtractArgs {
    const uint64_t* table_keys;
    const uint32_t* table_rep;
    const uint32_t* table_log2;
    uint32_t* rank_of_slot;
    const uint8_t* parent_masks;
    const uint32_t* parent_children;
    const uint32_t* child_rank_of_slot;  // non-null: parent_children hold child slots, written as ranks
    uint64_t* out_hashes;
    uint8_t* out_masks;
    uint32_t* out_children;
    uint32_t* tile_counts;    // tiles(slot_bound, kExtractTileItems)
    const uint64_t* offsets;
    uint64_t* slots_scratch;  // receives the table's slot count
    uint64_t* q_count;        // receives the level's unique node count
    uint64_t* w_count;        // receives the level's word count (atomic adds)
};
void launch_level_extract(hipStream_t s, const LevelExtractArgs& a, uint64_t slot_bound);

void launch_level_check(hipStream_t s, const uint32_t* parent_slots, const uint32_t* table_rep,
                        const uint8_t* parent_masks, const uint32_t* parent_children, const uint64_t* count_ptr,
                        uint64_t* collisions, uint64_t n_bound);

// ---- final layout ---------------------------------------------------------
struct FinalLevel {
    const uint8_t* masks;
    const uint32_t* children;
    uint64_t first_node;   // index of the level's first node in the concatenation
    uint64_t num_nodes;
};
struct FinalArgs {
    FinalLevel levels[SVDAG_MAX_LEVELS];
    int num_levels;
    uint64_t total_nodes;
    const uint64_t* leaves;
    uint64_t num_leaves;
    uint64_t leaves_start;   // word index of the first leaf
    uint32_t* positions;     // [total_nodes] absolute word offset of each node
    uint32_t* words;
    uint32_t* counts;        // tiles(total_nodes, kFinalTileItems)
    uint64_t* offsets;
};
void launch_final(hipStream_t s, const FinalArgs& a);

// ---- device voxeliser ------------------------------------------------------
// Per-triangle constants of the thin surface rule: normal n, n . v0, the
// projected edge functions (nx, ny, d) per axis plane and edge, the dominant
// axis a with the column axes u, w, the planes that are not edge-on, and the
// column rectangle (u0, w0, nw columns along w).
struct TriSetup {
    double n[3];
    double nd;
    double ex[3][3], ey[3][3], ed[3][3];
    int64_t u0, w0;
    uint32_t nw;
    int32_t a, u, w, planes;
};
constexpr uint64_t kVoxTileItems = 2048;  // columns per voxeliser tile
constexpr uint64_t kVoxStageMaxBytes = 4ull << 30;  // default largest one-pass staging area

// Per-row table of the triangles' column rectangles (a row runs along w).
struct VoxRows {
    uint32_t* tri;      // triangle of the row
    uint32_t* u;        // its u coordinate
    uint32_t* w0;       // first column kept
    uint32_t* cols;     // columns kept (a conservative superset of those the rule can emit)
    uint64_t* col_off;  // exclusive scan of cols
};

// rows[t] = rows of triangle t's rectangle (tris: 9 doubles per triangle);
// *rect_columns += the rectangles' columns (a bound on the kept columns); a
// non-finite coordinate sets kErrKeyRange in *err.
void launch_tri_setup(hipStream_t s, const double* tris, uint64_t n_tris, int depth, TriSetup* setup, uint32_t* rows,
                      uint64_t* rect_columns, unsigned long long* err);
// rows.tri / u / w0 / cols / col_off for the *n_rows_dev rows (n_rows on the
// host), the column total into *total_cols_dev; row_off = exclusive scan of
// the per-triangle row counts; tile_cols / tile_off: one entry per
// kVoxTileItems rows; a tile of rows with 2^32 or more columns sets
// kErrInternal in *err.
void launch_row_setup(hipStream_t s, const TriSetup* setup, const uint64_t* row_off, const uint64_t* n_tris_dev,
                      const uint64_t* n_rows_dev, uint64_t n_rows, const VoxRows& rows, uint32_t* tile_cols,
                      uint64_t* tile_off, uint64_t* total_cols_dev, unsigned long long* err);
// Voxels per tile of kVoxTileItems global columns (*total_cols_dev in all, at
// most total_cols_bound).  tile_row (tiles + 1 entries) receives each tile's
// first row; the write pass reads it.
void launch_vox_count(hipStream_t s, const TriSetup* setup, const VoxRows& rows, const uint64_t* n_rows_dev,
                      const uint64_t* total_cols_dev, uint64_t total_cols_bound, int depth, uint64_t* tile_row,
                      uint32_t* tile_counts);
void launch_vox_write(hipStream_t s, const TriSetup* setup, const VoxRows& rows, const uint64_t* tile_row,
                      const uint64_t* n_rows_dev, const uint64_t* total_cols_dev, uint64_t total_cols_bound, int depth,
                      const uint64_t* tile_offsets, uint64_t* keys);
// One evaluation of the columns: tile t's keys go to stage[t * 2 * kVoxTileItems
// ...] (2 * kVoxTileItems keys per tile of staging), their count to
// tile_counts[t]; total_cols = the columns (known on the host).  Then each
// tile's keys move to tile_offsets[t] (exclusive scan of the counts).
void launch_vox_emit(hipStream_t s, const TriSetup* setup, const VoxRows& rows, const uint64_t* n_rows_dev,
                     const uint64_t* total_cols_dev, uint64_t total_cols, int depth, uint64_t* tile_row,
                     uint32_t* tile_counts, uint64_t* stage);
void launch_vox_compact(hipStream_t s, const uint64_t* stage, const uint32_t* tile_counts, const uint64_t* tile_offsets,
                        const uint64_t* total_cols_dev, uint64_t total_cols, uint64_t* keys);

// Per pixel: the first voxel hit and its colour (the reference's trace_paths).
void launch_trace_paths(hipStream_t s, const uint32_t* words, uint64_t n_words, int depth,
                        const svdag_trace_params& p, uint32_t* paths, uint32_t* colors);

// out[i] = occupancy of voxel codes[i] in the flat layout (1 set, 0 not set,
// 2 the walk left the words).
void launch_query_voxels(hipStream_t s, const uint32_t* words, uint64_t n_words, int depth, const uint64_t* codes,
                         uint64_t n, uint8_t* out);

// ---- merge ----------------------------------------------------------------
// Remaps and inserts the nodes of many sub-DAG levels into one shared table.
struct MergeSource {
    const uint8_t* masks;
    const uint32_t* children;
    const uint64_t* hashes;
    uint64_t first;          // index of the source's first node in the concatenation
    uint64_t num;
    uint64_t child_first;    // index of the source's first child in the child concatenation
};
void launch_merge_insert(hipStream_t s, const MergeSource* d_sources, uint32_t num_sources, uint64_t total,
                         const uint32_t* child_slot_of, const uint32_t* child_rank_of_slot, uint8_t* masks_out,
                         uint32_t* children_out, uint64_t* table_keys, uint32_t* table_rep, const uint32_t* table_log2,
                         uint32_t* slot_of, BuildCounters* counters, int ordered, uint64_t* list_keys,
                         uint32_t* list_slots);

// ---- gathered chunks ------------------------------------------------------
// Source i's keys go to dst[start, start of source i + 1) (the last source up
// to `total`); starts strictly increase.
struct GatherSource {
    const uint64_t* keys;
    uint64_t start;
};
void launch_gather_keys(hipStream_t s, const GatherSource* src, uint32_t nsrc, uint64_t total, uint64_t* dst);

// ---- resident DAG (served mode) ------------------------------------------
// Nodes of one round level into the persistent table of that level.
struct ResidentLevelArgs {
    const uint64_t* hashes;       // the round level (ascending hash order)
    const uint8_t* masks;
    const uint32_t* children;     // round ranks of the level below
    uint64_t count;
    const uint32_t* child_slot_of;  // round rank of the level below -> resident slot
    uint64_t* table_keys;         // resident ordered table (2^lg + kTableOverflow slots)
    uint32_t* table_rep;
    uint32_t lg;
    uint8_t* rec_masks;           // slot-indexed records
    uint32_t* rec_children;
    uint32_t* slot_of;            // out: resident slot of each round node
    unsigned long long* occupied; // += keys placed
    unsigned long long* err;
};
void launch_resident_level(hipStream_t s, const ResidentLevelArgs& a);
void launch_resident_leaves(hipStream_t s, const uint64_t* leaves, uint64_t count, uint64_t* table_keys, uint32_t lg,
                            uint32_t* slot_of, unsigned long long* occupied, unsigned long long* err);

// Chunk roots of a round: its levels 0 .. top_levels - 1 lead to the chunk
// slots (8^top_levels of them, Morton order).
constexpr int kMaxResidentTop = kTailTopLevel + 1;
struct ResidentRootArgs {
    int top_levels;
    bool has_root;                               // the round has a root node
    const uint8_t* masks[kMaxResidentTop];       // round levels 0 .. top_levels - 1
    const uint32_t* children[kMaxResidentTop];
    const uint32_t* root_slot_of;                // round rank at level top_levels -> resident slot
    const uint64_t* root_table_keys;             // resident table of level top_levels
    uint32_t* root_slot;                         // per chunk slot: resident root slot (kEmptyChild: none)
    uint64_t* root_hash;                         // out: per chunk slot, its root's hash if added now, else 0
    unsigned long long* err;
};
void launch_resident_roots(hipStream_t s, const ResidentRootArgs& a);
// Filled chunk slots in slot order: codes[i] = slot index, list_slots[i] = root slot, *count.
void launch_resident_root_list(hipStream_t s, const uint32_t* root_slot, uint32_t slots, uint64_t* codes,
                               uint32_t* list_slots, uint64_t* count);

}  // namespace svdag
