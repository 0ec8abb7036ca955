/*
 * svdag.h -- C ABI of the MI355X (gfx950) sparse-voxel-DAG build.
 *
 * Drop-in boundary for the reference's DAG build API
 * (Phyronnaz/DAG_Compression_Fast, src/DAGCompression.h:42-50):
 *
 *   FCpuDag CreateDAG(const TStaticArray<FMortonCode, GPU>& Fragments, int32 Depth);
 *   FCpuDag MergeDAGs(std::vector<FCpuDag>&& CpuDags);
 *   TStaticArray<uint32, GPU> CreateFinalDAG(FCpuDag&& CpuDag);
 *   void FreeAll();
 *
 * Plain pointers and sizes only.  Every entry point returns 0 on success and a
 * non-zero svdag_status otherwise, with a message in svdag_last_error(); none
 * aborts (the reference prints and aborts, src/Core.h:111-128, 196-217).
 *
 * Output contract (bit-exact with the reference):
 *   levels[0] is the root level (one node for non-empty input), levels[depth-3]
 *   the level above the leaves (src/DAGCompression.cu:1008, 1077, 1093);
 *   nodes of a level are in ascending order of their 64-bit hash, leaves in
 *   ascending mask order (src/DAGCompression.cu:864-914);
 *   children[8*i + slot] indexes the next level (or the leaves), 0xFFFFFFFF
 *   marks an empty slot (src/DAGCompression.cu:952-957);
 *   hash = low word of MurmurHash128xN over the 8 child hashes, 0 -> 1
 *   (src/DAGCompression.cu:758-774, src/Utils.h:119-154).
 */
#ifndef SVDAG_H
#define SVDAG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SVDAG_MAX_LEVELS 24

typedef enum svdag_status {
    SVDAG_OK = 0,
    SVDAG_ERR_ARGUMENT = 1,   /* bad depth, null pointer, count, key >= 2^(3*depth) */
    SVDAG_ERR_HIP = 2,        /* a HIP runtime call failed                          */
    SVDAG_ERR_MEMORY = 3,     /* device allocation failed or exceeds the budget     */
    SVDAG_ERR_INTERNAL = 4,   /* a device-side invariant failed (e.g. table overflow)*/
    SVDAG_ERR_CAPACITY = 5    /* result does not fit the uint32 index space         */
} svdag_status;

typedef enum svdag_location {
    SVDAG_HOST = 0,           /* arrays in host memory (the reference's FCpuDag)    */
    SVDAG_DEVICE = 1          /* arrays in device memory (HBM)                      */
} svdag_location;

/* One level: mirror of FCpuLevel (src/DAGCompression.h:12-24). */
typedef struct svdag_level {
    uint8_t*  masks;          /* [num_nodes]                                        */
    uint32_t* children;       /* [num_nodes * 8]                                    */
    uint64_t* hashes;         /* [num_nodes], strictly ascending                    */
    uint64_t  num_nodes;
} svdag_level;

/* Mirror of FCpuDag (src/DAGCompression.h:26-40). */
typedef struct svdag_dag {
    int32_t     num_levels;               /* depth - 2                              */
    int32_t     location;                 /* svdag_location                         */
    svdag_level levels[SVDAG_MAX_LEVELS];
    uint64_t*   leaves;                   /* [num_leaves] 4x4x4 occupancy masks     */
    uint64_t    num_leaves;
    uint64_t    level_words[SVDAG_MAX_LEVELS]; /* sum of 1 + popc(mask) per level   */
    uint64_t    hash_collisions;          /* counted only when collision checks are on */
    /* Build statistics of create_dag (zero otherwise): */
    uint64_t    num_keys;                 /* N, input keys including duplicates     */
    uint64_t    leaf_positions;           /* L0, non-empty 4^3 blocks before dedup  */
    uint64_t    pre_dedup[SVDAG_MAX_LEVELS]; /* M_l, nodes of level l before dedup  */
    void*       storage;                  /* library-owned backing allocation       */
} svdag_dag;

typedef struct svdag_context svdag_context;

/* Options for svdag_set_option. */
typedef enum svdag_option {
    SVDAG_OPT_CHECK_COLLISIONS = 1,   /* 1: compare full node content on equal hash   */
    SVDAG_OPT_MEMORY_BUDGET = 2,      /* bytes of scratch the context may hold, 0 = no limit */
    SVDAG_OPT_TABLE_LOAD_SHIFT = 3,   /* dedup table slots >= 2^shift * nodes (1..3)  */
    SVDAG_OPT_HOST_THREADS = 4,       /* host threads for host-side work (merge)      */
    SVDAG_OPT_WAIT_SECONDS = 5,       /* completion deadline of a call (default 120 s); past it
                                         the call fails and names the stage that stalled */
    SVDAG_OPT_RADIX_BITS = 6,         /* widest digit of the key sort, 4..9 (default 9)      */
    SVDAG_OPT_SORT_TILE = 7,          /* keys per key-sort tile: 4096 or 8192 (default 4096) */
    SVDAG_OPT_TABLE_ESTIMATE = 8,     /* 1: size the leaf dedup table from a distinct-count
                                         estimate of the leaf masks (default); 0: from the leaf
                                         positions.  An estimate that proves too small makes the
                                         call rebuild with the bound-sized table. */
    SVDAG_OPT_VOXEL_STAGE_BYTES = 9,  /* largest staging area of svdag_voxelize_triangles
                                         (16 bytes per voxel column, default 4 GiB): within it
                                         the columns are evaluated once, past it twice (count,
                                         then write); 0 always takes two passes */
    SVDAG_OPT_SEG_FIRST_BITS = 10,    /* batched builds that sort inside each chunk (chunk
                                         depth <= 9): digit width of that sort's first pass,
                                         0 = half the chunk-local bits (default); clamped so
                                         that neither pass is wider than 9 bits */
    SVDAG_OPT_TOP_DIGIT_FIRST = 11    /* plain builds of 64-bit keys: sort the top digit first
                                         and the rest inside each top-digit bucket as 32-bit
                                         keys; 0 = never, 1 = from 2^20 keys on (default),
                                         2 = always.  Used where the bits below the top digit
                                         fit 32 and at most 12 of them need sorting. */
} svdag_option;

/* --- context ------------------------------------------------------------ */

/* A context owns a HIP stream binding, a device scratch arena and pinned
 * staging.  One context per host thread (the reference's global pool,
 * src/DAGCompression.cu:996-1001, is process-global and not thread-safe).
 * `stream` is a hipStream_t (NULL: the default stream). */
int svdag_context_create(svdag_context** ctx, void* stream);
int svdag_context_destroy(svdag_context* ctx);
int svdag_set_option(svdag_context* ctx, int32_t option, int64_t value);

/* Device scratch bytes create_dag needs for n keys at this depth (admission
 * input for serving). */
uint64_t svdag_scratch_bytes(uint64_t n, int32_t depth, int32_t key_bytes);

/* --- the build ------------------------------------------------------------ */

/* DAGCompression::CreateDAG (src/DAGCompression.cu:1003-1107).
 * d_keys: device pointer to n unsorted Morton codes with duplicates, each
 * < 2^(3*depth); 3 <= depth <= 21; n < 2^30.  The input is not modified
 * (the reference clobbers it, src/DAGCompression.cu:1028-1033).
 * On return `out` is a device-resident DAG owned by the caller
 * (svdag_free_dag); the call synchronises its stream once at the end. */
int svdag_create_dag(svdag_context* ctx, const uint64_t* d_keys, uint64_t n,
                     int32_t depth, svdag_dag* out);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (tight serving loops, harnesses that forbid hipMalloc, work set up for
 * graph capture): svdag_create_dag with its scratch in a fixed device array of
 * the library (SVDAG_NOALLOC_SCRATCH_BYTES, one copy per GPU, held until the
 * process ends) and its result in the caller's device buffer d_storage, of
 * storage_bytes >= svdag_result_bytes(n, depth).  Calls no hipMalloc, hipFree
 * or allocating routine.  A build needing more scratch than the array holds
 * (svdag_scratch_bytes(n, depth, 8) at load shift 1; about 5M keys at depth 9,
 * 1M at depth 10, 900k at depth 11, 650k at depth 12) or more storage than given fails with SVDAG_ERR_MEMORY
 * before any device work.  `out` points into d_storage and owns nothing
 * (out->storage is NULL; svdag_free_dag releases nothing).  Every call on a
 * GPU shares that GPU's array, which keeps the last call's contents (each call
 * clears what must start at zero): calls that overlap on one GPU, from any
 * context or stream, race and must be serialised by the caller. */
#define SVDAG_NOALLOC_SCRATCH_BYTES (256ull << 20)
int svdag_create_dag_noalloc(svdag_context* ctx, const uint64_t* d_keys, uint64_t n,
                             int32_t depth, void* d_storage, uint64_t storage_bytes,
                             svdag_dag* out);
/* Result bytes svdag_create_dag_noalloc needs for n keys at depth. */
uint64_t svdag_result_bytes(uint64_t n, int32_t depth);

/* Same with 32-bit keys (depth <= 10): the 512^3 u32 variant. */
int svdag_create_dag_u32(svdag_context* ctx, const uint32_t* d_keys, uint64_t n,
                         int32_t depth, svdag_dag* out);

/* DAGCompression::MergeDAGs (src/DAGCompression.cu:1109-1320, 1447-1548).
 * `count` DAGs of equal depth, a power of 8, in Morton order of their
 * sub-volumes (src/Utils/Aabb.h:83-105); host- or device-resident.  Inputs are
 * not consumed.  Output: device-resident DAG with log8(count) more levels. */
int svdag_merge_dags(svdag_context* ctx, const svdag_dag* dags, uint64_t count,
                     svdag_dag* out);

/* Batched build: MergeDAGs(CreateDAG(chunk_i)) for `num_chunks` (a power of
 * 8) chunks in one device pass sharing one dedup table per level.
 * d_keys holds the chunks back to back, chunk i at [offsets[i], offsets[i+1])
 * (offsets: host array of num_chunks + 1), keys chunk-local (< 2^(3*chunk_depth)).
 * Output equals svdag_merge_dags of the per-chunk svdag_create_dag results. */
int svdag_create_dag_batched(svdag_context* ctx, const uint64_t* d_keys,
                             const uint64_t* offsets, uint64_t num_chunks,
                             int32_t chunk_depth, svdag_dag* out);

/* The same batched build with each chunk's keys in its own device array (the
 * reference's per-chunk flow, where every chunk's fragments sit in the
 * voxelizer's GPU buffer, src/DAG_Compression_Fast.cpp:283-302):
 * d_chunk_keys[i] (host array of device pointers) holds counts[i] keys of
 * chunk i; chunks with no keys may have a null pointer.  The keys are gathered
 * back to back on the device first (8 bytes per key of extra memory). */
int svdag_create_dag_batched_gather(svdag_context* ctx, const uint64_t* const* d_chunk_keys,
                                    const uint64_t* counts, uint64_t num_chunks,
                                    int32_t chunk_depth, svdag_dag* out);

/* The batched build captured once as a HIP graph and replayed.  No
 * counterpart in the reference, whose batch loop synchronises the host per
 * chunk and per level (src/DAG_Compression_Fast.cpp:283-302,
 * src/DAGCompression.cu:252-268).  The graph reads the keys at d_keys and the
 * num_chunks + 1 offsets at d_offsets (both device buffers the caller keeps and
 * may rewrite between launches) and holds its scratch and result for up to
 * max_keys keys (< 2^30).  Offsets that do not start at 0, decrease, or end
 * past max_keys make that launch's result SVDAG_ERR_ARGUMENT.  The graph has
 * its own stream: svdag_batched_graph_launch enqueues a replay without
 * waiting; svdag_batched_graph_result waits for it and returns the DAG, whose
 * arrays live in the graph's storage until its next launch (out->storage is
 * NULL). */
typedef struct svdag_batched_graph svdag_batched_graph;
int svdag_batched_graph_create(svdag_context* ctx, const uint64_t* d_keys, const uint64_t* d_offsets,
                               uint64_t num_chunks, int32_t chunk_depth, uint64_t max_keys,
                               svdag_batched_graph** out);
int svdag_batched_graph_launch(svdag_batched_graph* g);
int svdag_batched_graph_result(svdag_batched_graph* g, svdag_dag* out);
int svdag_batched_graph_destroy(svdag_batched_graph* g);

/* DAGCompression::CreateFinalDAG (src/DAGCompression.cu:1550-1610): the flat
 * uint32 layout.  *d_words is a device allocation (svdag_free_words). */
int svdag_create_final_dag(svdag_context* ctx, const svdag_dag* dag,
                           uint32_t** d_words, uint64_t* n_words);

/* Copies between host (FCpuDag-like) and device residency.  `out` is a new
 * DAG owned by the caller. */
int svdag_dag_to_host(svdag_context* ctx, const svdag_dag* dag, svdag_dag* out);
int svdag_dag_to_device(svdag_context* ctx, const svdag_dag* dag, svdag_dag* out);

int svdag_free_dag(svdag_context* ctx, svdag_dag* dag);
int svdag_free_words(svdag_context* ctx, uint32_t* d_words);

/* DAGCompression::FreeAll (src/DAGCompression.cu:1612-1615): release the
 * context's scratch arena and its cached result blocks. */
int svdag_free_all(svdag_context* ctx);

/* Device memory a context holds: its scratch arena, result blocks owned by
 * live DAGs and word arrays, and freed result blocks kept for reuse (these
 * are released, and the allocation retried once, when an allocation fails or
 * would pass SVDAG_OPT_MEMORY_BUDGET: the reference's out-of-memory fallback,
 * src/Memory.cpp:56-70). */
int svdag_memory_usage(svdag_context* ctx, uint64_t* scratch_bytes, uint64_t* result_bytes,
                       uint64_t* cached_bytes);

/* --- resident global DAG (served mode) ----------------------------------- */

/* No single counterpart in the reference: the global DAG of the reference's
 * chunk loop (src/DAG_Compression_Fast.cpp:283-320, merged by
 * DAGCompression::MergeDAGs, src/DAGCompression.cu:1109-1320, 1447-1548), kept
 * on the device while chunks are added to it.  The grid holds 8^slot_levels
 * chunk slots (Morton order) of side 2^chunk_depth; the global depth is
 * chunk_depth + slot_levels (1 <= slot_levels <= 4).  Each level keeps one
 * dedup table for the object's life, sized for min(max_nodes, 8^l) distinct
 * nodes at global level l (leaves: min(max_nodes, 8^(depth-2))); a level can
 * never hold more nodes than the keys of all chunks added, and holds far fewer
 * on surface-like data (about 1.5% of the keys at the 512-chunk workload).
 * Device memory: about 90 bytes per node of capacity per level.  Not
 * thread-safe; it runs on its context's stream. */
typedef struct svdag_resident svdag_resident;
int svdag_resident_create(svdag_context* ctx, int32_t chunk_depth, int32_t slot_levels,
                          uint64_t max_nodes, svdag_resident** out);
int svdag_resident_destroy(svdag_resident* r);
/* Adds the chunks of `round`, a device DAG of depth chunk_depth + slot_levels
 * (as svdag_create_dag_batched returns for the 8^slot_levels chunk slots, empty
 * ones allowed).  Every chunk of the round must be at a slot not filled yet.
 * root_hashes (host, 8^slot_levels entries, or NULL) receives, per slot, the
 * root hash of the chunk this call added there (its content identity), else 0.
 * A round whose nodes could overflow a table fails with SVDAG_ERR_MEMORY before
 * any device work; a failure after device work has started leaves the object
 * unusable. */
int svdag_resident_insert(svdag_resident* r, const svdag_dag* round, uint64_t* root_hashes);
/* Nodes per resident level (global levels slot_levels .. depth-3, then the
 * leaves: chunk_depth - 1 entries) and the number of chunks added. */
int svdag_resident_counts(svdag_resident* r, uint64_t* nodes, uint64_t* chunks);
/* The global DAG of every chunk added so far, in the reference's form (each
 * level in ascending hash order): a new device DAG owned by the caller, equal
 * to svdag_merge_dags of the chunks' DAGs (empty slots empty). */
int svdag_resident_dag(svdag_resident* r, svdag_dag* out);

/* Message of the last failing call on this host thread. */
const char* svdag_last_error(void);

/* --- the DAG file and point queries --------------------------------------- */

/* The reference's on-disk DAG (FFileWriter / FFileReader framing,
 * src/Serializer.h:10-101, read at src/DAG_Compression_Fast.cpp:361-375): six
 * float64 (AABB min x, y, z, max x, y, z), uint32 level count, uint64 word
 * count, the uint32 words of the flat layout.  Native little-endian, no padding.
 * words_location: SVDAG_HOST, or SVDAG_DEVICE (copied out on ctx's stream;
 * ctx may be NULL for host words). */
int svdag_write_dag_file(svdag_context* ctx, const char* path, const double* aabb_min,
                         const double* aabb_max, uint32_t levels, const uint32_t* words,
                         uint64_t n_words, int32_t words_location);
/* Reads such a file; *words is malloc'd host memory (svdag_gen_free).  A file
 * shorter than its word count says is an error. */
int svdag_read_dag_file(const char* path, double* aabb_min, double* aabb_max, uint32_t* levels,
                        uint32_t** words, uint64_t* n_words);

/* Device fragment producer: voxelises a triangle soup into unsorted Morton
 * keys with duplicates (the reference's input contract; it rasterises with
 * OpenGL, src/Voxelizer.cpp:254-369).  d_tris: device, 9 float64 per triangle
 * (x y z of its three corners) in a 2^depth grid; the thin surface rule of the
 * triangle-surface generator, evaluated identically (no fused multiply-adds),
 * so the voxel set equals svdag_gen_triangle_surfaces' for the same triangles.
 * *d_keys is a device allocation (svdag_free_keys); 1 <= depth <= 21.  A
 * non-finite coordinate is SVDAG_ERR_ARGUMENT. */
int svdag_voxelize_triangles(svdag_context* ctx, const double* d_tris, uint64_t n_tris,
                             int32_t depth, uint64_t** d_keys, uint64_t* n_keys);
int svdag_free_keys(svdag_context* ctx, uint64_t* d_keys);

/* The reference's SortKeys (src/DAGCompression.cu:336-357; cub::DeviceRadixSort::
 * SortKeys on bits [0, end_bit)): d_out = the n keys of d_keys in ascending order
 * of their low `bits` bits, stable as CUB's: keys equal in those bits keep their
 * input order.  Device pointers; d_out must not alias d_keys, which is left
 * untouched; 1 <= bits <= 64, n < 2^32.  Scratch: 16 bytes per key plus the
 * digit counts, in the context's arena. */
int svdag_sort_keys(svdag_context* ctx, const uint64_t* d_keys, uint64_t n, int32_t bits,
                    uint64_t* d_out);
/* The same for 32-bit keys (1 <= bits <= 32; 8 bytes of scratch per key): the
 * reference's 32-bit Morton variant (MORTON_BITS, src/Voxelizer.cpp:207). */
int svdag_sort_keys_u32(svdag_context* ctx, const uint32_t* d_keys, uint64_t n, int32_t bits,
                        uint32_t* d_out);

/* Occupancy of voxels in a flat layout of a depth-`depth` DAG, walked as the
 * reference's tracer reads it (src/DAGTracer.cu:85-176).  d_words, d_codes and
 * d_out are device pointers; d_out[i] = 1 if Morton code d_codes[i] is set, 0 if
 * not (or the code is >= 2^(3*depth)), 2 if the walk leaves the words. */
int svdag_query_voxels(svdag_context* ctx, const uint32_t* d_words, uint64_t n_words,
                       int32_t depth, const uint64_t* d_codes, uint64_t n, uint8_t* d_out);

/* Pinhole camera of the tracer, in grid units (the reference's
 * FDAGTracer::TracePathsParams, src/DAGTracer.h:17-29): pixel (x, y) looks
 * along ray_min + x * ray_ddx + y * ray_ddy - camera_position. */
typedef struct svdag_trace_params {
    uint32_t debug_level;  /* 0xFFFFFFFF: colour = low bytes of the hit position;
                            * else MurmurHash32 of the word index of the node at
                            * level depth - 3 - debug_level on the path */
    uint32_t width;
    uint32_t height;
    uint32_t reserved;     /* 0 */
    double camera_position[3];
    double ray_min[3];
    double ray_ddx[3];
    double ray_ddy[3];
} svdag_trace_params;

/* The reference's FDAGTracer::GetTraceParams (src/DAGTracer.h:33-82): camera at
 * `position` with the rows right, up, direction of its rotation, full vertical
 * field of view `fov_degrees`, mapped into the 2^depth grid of the box
 * [aabb_min, aabb_max]. */
int svdag_trace_params_from_view(const double* position, const double* right, const double* up,
                                 const double* direction, double fov_degrees, uint32_t width,
                                 uint32_t height, const double* aabb_min, const double* aabb_max,
                                 int32_t depth, uint32_t debug_level, svdag_trace_params* out);

/* Primary rays through a flat layout (the reference's trace_paths kernel,
 * src/DAGTracer.cu:177-428, launched by FDAGTracer::ResolvePaths, :430-444).
 * Per pixel (row-major, y * width + x): d_paths[3 * pixel + 0..2] = grid x, y,
 * z of the first voxel hit, (0, 0, 0) when the ray hits nothing; d_colors[pixel]
 * = the reference's colour (rgb888 of the position's low bytes, or the debug
 * hash).  A walk that leaves the words gives (0xFFFFFFFF x 3) and colour 0.
 * Either output may be NULL; all pointers are device pointers; 3 <= depth <= 21. */
int svdag_trace_paths(svdag_context* ctx, const uint32_t* d_words, uint64_t n_words,
                      int32_t depth, const svdag_trace_params* params, uint32_t* d_paths,
                      uint32_t* d_colors);

/* --- per-kernel timing ---------------------------------------------------- */

/* Counterpart of the reference's per-operation timer FCudaScopePerfRecorder
 * (src/CudaPerfRecorder.h:5-60, used at src/DAGCompression.cu:352, 384, 413 ...),
 * which synchronises after each operation.  Here bit i of kernel_mask brackets
 * every launch of kernel i (svdag_kernel_name(i)) with two HIP events on the
 * context's stream; the pairs are resolved at the call's own synchronisation and
 * summed per kernel.  0 switches timing off. */
int svdag_profile_enable(svdag_context* ctx, uint64_t kernel_mask);
int svdag_profile_reset(svdag_context* ctx);
int svdag_profile_read(svdag_context* ctx, int32_t kernel, double* total_ms, uint64_t* launches);
/* Name of kernel i, NULL past the last one. */
const char* svdag_kernel_name(int32_t kernel);

/* --- synthetic inputs (host) --------------------------------------------- */

/* Morton code of (x, y, z) (src/DAG_Compression_Fast.cpp:223-236). */
uint64_t svdag_morton_encode(uint32_t x, uint32_t y, uint32_t z);

/* Seeded generators; *keys is malloc'd (svdag_gen_free). */
int svdag_gen_sphere_shells(uint64_t seed, int32_t depth, int32_t count, double rmin,
                            double rmax, double dup_fraction, int32_t threads,
                            uint64_t** keys, uint64_t* n);
int svdag_gen_triangle_surfaces(uint64_t seed, int32_t depth, int32_t count, double mu,
                                double sigma, double dup_fraction, int32_t threads,
                                uint64_t** keys, uint64_t* n);
/* The triangles of svdag_gen_triangle_surfaces(seed, depth, count, mu, sigma,
 * ...): out receives 9 float64 per triangle (corners v0, v1, v2). */
int svdag_gen_triangles(uint64_t seed, int32_t depth, int32_t count, double mu, double sigma,
                        double* out);
/* Host voxeliser of the triangle-surface rule (the same rule as
 * svdag_voxelize_triangles): sorted unique Morton keys of host triangles
 * (9 float64 each); *keys is malloc'd (svdag_gen_free). */
int svdag_gen_voxelize(const double* tris, uint64_t n_tris, int32_t depth, int32_t threads,
                       uint64_t** keys, uint64_t* n);
int svdag_gen_value_noise(uint64_t seed, int32_t depth, int32_t spacing, uint32_t sub_depth,
                          uint64_t sub_index, int32_t threads, uint64_t** keys, uint64_t* n);
int svdag_gen_chunk_plan(uint64_t seed, uint64_t num_chunks, double median, double sigma,
                         uint64_t nmin, uint64_t nmax, uint64_t* n_out,
                         uint64_t* unique_out, int32_t* kind_out);
int svdag_gen_chunk(uint64_t seed, uint64_t chunk, int32_t depth, double median, double sigma,
                    uint64_t nmin, uint64_t nmax, uint64_t* keys, uint64_t capacity,
                    uint64_t* n);
void svdag_gen_free(void* p);

#ifdef __cplusplus
}
#endif

#endif /* SVDAG_H */
