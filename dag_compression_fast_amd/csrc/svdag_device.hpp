// svdag_device.hpp -- device helpers shared by the gfx950 kernels.
//
// Wave64 scans and reductions, agent-scope atomics, and the node hash (MurmurHash128xN of the reference,
// src/Utils.h:119-154, as used by HashChildrenHashes, src/DAGCompression.cu:758-763).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace svdag {

constexpr uint32_t kEmptyChild = 0xFFFFFFFFu;
constexpr int kWave = 64;

// s_waitcnt operand (gfx9 encoding) that waits for vmcnt 0 and leaves expcnt
// and lgkmcnt at their maxima.  Placed where a loop consumes prefetched loads,
// so that no load is pending at the loop head.
constexpr int kWaitVmcnt0 = 0x0F70;

__device__ __forceinline__ uint64_t atomic_load_relaxed(const uint64_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t atomic_load_relaxed(const uint32_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Probe read of a dedup-table slot.  Slots only ever change from 0 to their
// final key, so a value read through the CU's L1 is either final or a stale 0,
// which the caller settles with a CAS; hot slots then stay in L1.
__device__ __forceinline__ uint64_t probe_load(const uint64_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void atomic_store_relaxed(uint64_t* p, uint64_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void atomic_store_relaxed(uint32_t* p, uint32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ uint64_t lanemask_lt() { return (1ull << lane_id()) - 1ull; }

// ---------------------------------------------------------------------------
// Hash: Utils::Rotl64 (Utils.h:41-44), Utils::MurmurHash64 (Utils.h:86-94),
// Utils::MurmurHash128xN with Size = 4, Seed = 0 (Utils.h:119-154).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t rotl64(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }

__device__ __forceinline__ uint64_t fmix64(uint64_t h)
{
    h ^= h >> 33;
    h *= 0xff51afd7ed558ccdull;
    h ^= h >> 33;
    h *= 0xc4ceb9fe1a85ec53ull;
    h ^= h >> 33;
    return h;
}

// Low word of MurmurHash128xN(child hashes as 4 x uint128, 4, 0); 0 -> 1.
__device__ __forceinline__ uint64_t hash_children(const uint64_t ch[8])
{
    const uint64_t c1 = 0x87c37b91114253d5ull;
    const uint64_t c2 = 0x4cf5ad432745937full;
    uint64_t h1 = 0, h2 = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        uint64_t k1 = ch[2 * i + 0];
        uint64_t k2 = ch[2 * i + 1];
        k1 *= c1; k1 = rotl64(k1, 31); k1 *= c2; h1 ^= k1;
        h1 = rotl64(h1, 27); h1 += h2; h1 = h1 * 5 + 0x52dce729;
        k2 *= c2; k2 = rotl64(k2, 33); k2 *= c1; h2 ^= k2;
        h2 = rotl64(h2, 31); h2 += h1; h2 = h2 * 5 + 0x38495ab5;
    }
    h1 ^= 4;  // the reference XORs the block count, not the byte length
    h2 ^= 4;
    h1 += h2;
    h2 += h1;
    h1 = fmix64(h1);
    h2 = fmix64(h2);
    h1 += h2;
    return h1 == 0 ? 1 : h1;
}

// ---------------------------------------------------------------------------
// Wave64 / block scans.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v)
{
    const uint32_t lane = lane_id();
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const uint32_t t = __shfl_up(v, d, kWave);
        if (lane >= uint32_t(d)) v += t;
    }
    return v;
}

__device__ __forceinline__ uint64_t wave_reduce_sum(uint64_t v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, kWave);
    return v;
}

__device__ __forceinline__ uint32_t wave_reduce_sum32(uint32_t v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, kWave);
    return v;
}

// Exclusive block scan of one value per thread.  `scratch` holds THREADS/64
// words of LDS.  Contains two barriers; every thread of the block must call it.
template <int THREADS>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* scratch, uint32_t& total)
{
    constexpr int W = THREADS / kWave;
    const int w = threadIdx.x / kWave;
    const uint32_t inc = wave_inclusive_scan(v);
    if (lane_id() == kWave - 1) scratch[w] = inc;
    __syncthreads();
    uint32_t base = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < W; ++i) {
        const uint32_t x = scratch[i];
        base += i < w ? x : 0;
        tot += x;
    }
    __syncthreads();
    total = tot;
    return base + inc - v;
}

// Block sum (two barriers; every thread must call it).
template <int THREADS>
__device__ __forceinline__ uint64_t block_reduce_sum(uint64_t v, uint64_t* scratch)
{
    constexpr int W = THREADS / kWave;
    v = wave_reduce_sum(v);
    if (lane_id() == 0) scratch[threadIdx.x / kWave] = v;
    __syncthreads();
    uint64_t tot = 0;
#pragma unroll
    for (int i = 0; i < W; ++i) tot += scratch[i];
    __syncthreads();
    return tot;
}

// 64-bit exclusive block scan (two barriers; every thread must call it).
template <int THREADS>
__device__ __forceinline__ uint64_t block_exclusive_scan64(uint64_t v, uint64_t* scratch, uint64_t& total)
{
    constexpr int W = THREADS / kWave;
    const int w = threadIdx.x / kWave;
    const uint32_t lane = lane_id();
    uint64_t inc = v;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const uint64_t t = __shfl_up(inc, d, kWave);
        if (lane >= uint32_t(d)) inc += t;
    }
    if (lane == kWave - 1) scratch[w] = inc;
    __syncthreads();
    uint64_t base = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < W; ++i) {
        const uint64_t x = scratch[i];
        base += i < w ? x : 0;
        tot += x;
    }
    __syncthreads();
    total = tot;
    return base + inc - v;
}

__host__ __device__ __forceinline__ uint32_t ceil_log2_u64(uint64_t v)
{
    uint32_t r = 0;
    while ((1ull << r) < v) ++r;
    return r;
}

}  // namespace svdag
