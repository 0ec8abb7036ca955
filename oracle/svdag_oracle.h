/*
 * svdag_oracle.h -- CPU oracle for the sparse-voxel-DAG build.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain C++ restatement of the reference
 * algorithm (Phyronnaz/DAG_Compression_Fast, src/DAGCompression.cu) used as the
 * checker for the HIP library and as the timed host-core baseline.  Only the
 * test suite, the smoke entry and the benchmark's CPU-baseline leg may load it;
 * the product library never links, calls or falls back to it.
 *
 * Every entry point cites the reference code it restates.  Parity pins (what
 * checks this oracle independently of itself) are listed per function below.
 */
#ifndef SVDAG_ORACLE_H
#define SVDAG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One interior level, mirror of FCpuLevel (src/DAGCompression.h:12-24). */
typedef struct ora_level {
    uint8_t*  masks;     /* [num_nodes] child masks                        */
    uint32_t* children;  /* [num_nodes * 8], 0xFFFFFFFF marks an empty slot */
    uint64_t* hashes;    /* [num_nodes], strictly ascending                */
    uint64_t  num_nodes;
} ora_level;

/* Mirror of FCpuDag (src/DAGCompression.h:26-40). levels[0] is the root level. */
typedef struct ora_dag {
    int32_t    num_levels;      /* depth - 2 */
    ora_level* levels;
    uint64_t*  leaves;          /* 4x4x4 occupancy masks, strictly ascending */
    uint64_t   num_leaves;
    uint64_t   hash_collisions; /* equal 64-bit hash, different content: merged as the reference does */
    /* Build statistics (create_dag only; zero after merge): */
    uint64_t   unique_keys;     /* U: distinct voxels                         */
    uint64_t   leaf_positions;  /* L0: non-empty 4^3 blocks before leaf dedup */
    uint64_t   pre_dedup[32];   /* M_l: Morton-distinct nodes of level l before dedup */
} ora_dag;

/* MurmurHash128xN (src/Utils.h:119-154) with the final length XOR exposed.
 * The reference XORs `size` (number of 16-byte blocks); canonical
 * MurmurHash3_x64_128 XORs the byte length.  Pinned against the published
 * SMHasher verification value of MurmurHash3_x64_128 (len_xor = 16*size). */
void ora_murmur128xN(const uint64_t* data, uint32_t size, uint32_t seed,
                     uint64_t len_xor, uint64_t out[2]);

/* HashChildrenHashes (src/DAGCompression.cu:758-763): low word of
 * MurmurHash128xN(child_hashes as 4 x uint128, 4, 0), 0 mapped to 1. */
uint64_t ora_hash_children(const uint64_t child_hashes[8]);

/* DAGCompression::CreateDAG (src/DAGCompression.cu:1003-1107).
 * keys: unsorted Morton codes with duplicates, each < 2^(3*depth).
 * 3 <= depth <= 21.  threads: host threads (>= 1).
 * Returns 0 on success, non-zero with ora_last_error() set otherwise. */
int ora_create_dag(const uint64_t* keys, uint64_t n, int32_t depth,
                   int32_t threads, ora_dag* out);

/* Same as ora_create_dag for 32-bit keys (depth <= 10). */
int ora_create_dag_u32(const uint32_t* keys, uint64_t n, int32_t depth,
                       int32_t threads, ora_dag* out);

/* DAGCompression::MergeDAGs (src/DAGCompression.cu:1109-1320, 1447-1548).
 * count must be a power of 8; dags are in Morton order of their sub-volumes
 * (src/Utils/Aabb.h:83-105).  Inputs are left untouched. */
int ora_merge_dags(const ora_dag* dags, uint64_t count, int32_t threads,
                   ora_dag* out);

/* DAGCompression::CreateFinalDAG (src/DAGCompression.cu:1550-1610).
 * *words is malloc'd; free with ora_free. */
int ora_create_final_dag(const ora_dag* dag, uint32_t** words, uint64_t* n_words);

/* Camera of the tracer in grid units, field for field the layout of
 * svdag_trace_params (src/DAGTracer.h:17-29). */
typedef struct ora_trace_params {
    uint32_t debug_level, width, height, reserved;
    double camera_position[3], ray_min[3], ray_ddx[3], ray_ddy[3];
} ora_trace_params;

/* trace_paths (src/DAGTracer.cu:177-428) over a flat layout: per pixel
 * (row-major) the grid position of the first voxel hit ((0,0,0) for none,
 * 0xFFFFFFFF x 3 when the walk leaves the words) and the reference's colour.
 * Parity pins: every hit is a voxel of the decoded set, and it is the voxel
 * the ray enters first up to float32 rounding, by brute force over all voxels
 * (tests/test_trace.py).  Either output may be NULL.  Returns non-zero on bad
 * arguments. */
int ora_trace_paths(const uint32_t* words, uint64_t n_words, int32_t depth,
                    const ora_trace_params* p, int32_t threads, uint32_t* paths,
                    uint32_t* colors);

void ora_free_dag(ora_dag* dag);
void ora_free(void* p);
const char* ora_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* SVDAG_ORACLE_H */
