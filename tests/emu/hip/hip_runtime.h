// Mock <hip/hip_runtime.h> for the CPU kernel emulator (tests only).
//
// Lets the product's .hip sources compile with g++: every GPU thread of a block
// runs as a cooperative fiber; __syncthreads, wave ballots and shuffles are
// rendezvous points.  Blocks run concurrently on a pool of OS threads (LDS is
// thread_local, atomics are real), or, under emu_set_schedule(1), one at a time
// in decreasing block order, so a block that depends on another block of the
// same launch shows up as a stall or a wrong result.  Launch configurations are
// checked (a zero or oversized grid or block is an error of hipGetLastError).
// Device memory is host memory filled with a garbage pattern, so code that
// relies on zeroed allocations misbehaves here too.
#pragma once

#include <cstddef>
#include <cstdint>
#include <cmath>
#include <cstring>
#include <functional>

#define __global__
#define __device__
#define __align__(n) __attribute__((aligned(n)))
#define HIP_SYMBOL(x) (&(x))
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static thread_local

struct dim3 {
    unsigned x = 1, y = 1, z = 1;
    dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct uint4 { unsigned x, y, z, w; };
inline uint4 make_uint4(unsigned a, unsigned b, unsigned c, unsigned d) { return uint4{a, b, c, d}; }

namespace emu {
const dim3& thread_idx();
const dim3& block_idx();
const dim3& block_dim();
const dim3& grid_dim();
void sync_block();
void yield_sleep();
unsigned lane();
unsigned long long ballot(bool p);
void wave_sync();
unsigned long long exchange(unsigned long long v, int src_lane);  // value of lane src (or own if out of range)
void launch(dim3 grid, dim3 block, std::function<void()> body);
}  // namespace emu

#define threadIdx (emu::thread_idx())
#define blockIdx (emu::block_idx())
#define blockDim (emu::block_dim())
#define gridDim (emu::grid_dim())

inline void __syncthreads() { emu::sync_block(); }
inline unsigned __lane_id() { return emu::lane(); }
inline unsigned long long __ballot(int p) { return emu::ballot(p != 0); }
inline int __popc(unsigned v) { return __builtin_popcount(v); }
inline int __popcll(unsigned long long v) { return __builtin_popcountll(v); }
#define __builtin_amdgcn_s_sleep(x) emu::yield_sleep()
#define __builtin_amdgcn_wave_barrier() emu::wave_sync()
#define __builtin_amdgcn_s_waitcnt(x) ((void)(x))
#define __builtin_amdgcn_sched_barrier(x) ((void)(x))
#define __builtin_nontemporal_load(p) (*(p))
using std::isfinite;

template <class T>
T __shfl_up(T v, unsigned d, int = 64)
{
    unsigned long long bits = 0;
    std::memcpy(&bits, &v, sizeof(T));
    const int src = int(emu::lane()) - int(d);
    bits = emu::exchange(bits, src);
    T out;
    std::memcpy(&out, &bits, sizeof(T));
    return out;
}
template <class T>
T __shfl_down(T v, unsigned d, int = 64)
{
    unsigned long long bits = 0;
    std::memcpy(&bits, &v, sizeof(T));
    bits = emu::exchange(bits, int(emu::lane() + d));
    T out;
    std::memcpy(&out, &bits, sizeof(T));
    return out;
}
template <class T>
T __shfl(T v, int src, int = 64)
{
    unsigned long long bits = 0;
    std::memcpy(&bits, &v, sizeof(T));
    bits = emu::exchange(bits, src & 63);
    T out;
    std::memcpy(&out, &bits, sizeof(T));
    return out;
}
// Lane reads of wave-uniform calls (every lane of the wave takes part here).
inline int emu_readlane(int v, int src) { return int(emu::exchange((unsigned long long)(unsigned)v, src & 63)); }
#define __builtin_amdgcn_readfirstlane(x) emu_readlane(int(x), 0)
#define __builtin_amdgcn_readlane(x, l) emu_readlane(int(x), int(l))
inline int __ffsll(long long v) { return __builtin_ffsll(v); }
inline int __clzll(long long v) { return v ? __builtin_clzll((unsigned long long)v) : 64; }
template <class T>
T __shfl_xor(T v, int m, int = 64)
{
    unsigned long long bits = 0;
    std::memcpy(&bits, &v, sizeof(T));
    bits = emu::exchange(bits, int(emu::lane() ^ unsigned(m)));
    T out;
    std::memcpy(&out, &bits, sizeof(T));
    return out;
}

#ifndef __HIP_MEMORY_SCOPE_WORKGROUP
#define __HIP_MEMORY_SCOPE_WORKGROUP 2
#endif
#ifndef __HIP_MEMORY_SCOPE_AGENT
#define __HIP_MEMORY_SCOPE_AGENT 3
#endif
#ifndef __clang__  // clang provides these as builtins
template <class T>
T __hip_atomic_load(const T* p, int, int) { return __atomic_load_n(p, __ATOMIC_RELAXED); }
template <class T, class U>
void __hip_atomic_store(T* p, U v, int, int) { __atomic_store_n(p, T(v), __ATOMIC_RELAXED); }
#endif

template <class T, class U>
T atomicAdd(T* p, U v) { return __atomic_fetch_add(p, T(v), __ATOMIC_SEQ_CST); }
template <class T, class U>
T atomicOr(T* p, U v) { return __atomic_fetch_or(p, T(v), __ATOMIC_SEQ_CST); }
template <class T, class U>
T atomicMax(T* p, U v)
{
    T old = __atomic_load_n(p, __ATOMIC_RELAXED);
    while (old < T(v) && !__atomic_compare_exchange_n(p, &old, T(v), false, __ATOMIC_SEQ_CST, __ATOMIC_RELAXED)) {
    }
    return old;
}
template <class T, class U, class V>
T atomicCAS(T* p, U cmp, V val)
{
    T expected = T(cmp);
    __atomic_compare_exchange_n(p, &expected, T(val), false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST);
    return expected;
}

template <class T>
inline T min(T a, T b) { return a < b ? a : b; }
template <class T>
inline T max(T a, T b) { return a > b ? a : b; }

// ---- host API ---------------------------------------------------------------
typedef int hipError_t;
enum { hipSuccess = 0, hipErrorOutOfMemory = 2, hipErrorInvalidValue = 1 };
typedef void* hipStream_t;
typedef void* hipEvent_t;
typedef void* hipMemPool_t;
enum hipMemcpyKind { hipMemcpyHostToHost = 0, hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2,
                     hipMemcpyDeviceToDevice = 3, hipMemcpyDefault = 4 };
enum { hipMemPoolAttrReleaseThreshold = 4 };
enum { hipHostMallocDefault = 0 };
struct hipDeviceProp_t { int multiProcessorCount; char name[256]; };

extern "C" {
hipError_t hipMalloc(void** p, size_t n);
hipError_t hipFree(void* p);
hipError_t hipGetSymbolAddress(void** p, const void* symbol);
hipError_t hipMallocAsync(void** p, size_t n, hipStream_t s);
hipError_t hipFreeAsync(void* p, hipStream_t s);
hipError_t hipHostMalloc(void** p, size_t n, unsigned flags);
hipError_t hipHostFree(void* p);
hipError_t hipMemcpy(void* d, const void* s, size_t n, int kind);
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, int kind, hipStream_t st);
hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t st);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipDeviceSynchronize();
hipError_t hipGetLastError();
const char* hipGetErrorString(hipError_t e);
hipError_t hipGetDevice(int* d);
hipError_t hipSetDevice(int d);
hipError_t hipGetDeviceCount(int* n);
hipError_t hipGetDeviceProperties(hipDeviceProp_t* p, int d);
hipError_t hipDeviceGetDefaultMemPool(hipMemPool_t* p, int d);
hipError_t hipMemPoolSetAttribute(hipMemPool_t p, int attr, void* v);
hipError_t hipMemPoolTrimTo(hipMemPool_t p, size_t n);
hipError_t hipStreamQuery(hipStream_t s);
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned flags);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s);
hipError_t hipEventQuery(hipEvent_t e);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventCreate(hipEvent_t* e);
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b);
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned flags);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned flags);
// Graphs: launches and copies issued while a stream is captured are recorded
// instead of run; a graph launch runs the recorded operations in order.
typedef void* hipGraph_t;
typedef void* hipGraphExec_t;
typedef void* hipGraphNode_t;
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal = 0, hipStreamCaptureModeThreadLocal = 1,
                            hipStreamCaptureModeRelaxed = 2 };
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode mode);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g);
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, hipGraphNode_t* err_node, char* log, size_t log_size);
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s);
hipError_t hipGraphExecDestroy(hipGraphExec_t e);
hipError_t hipGraphDestroy(hipGraph_t g);
}
enum { hipErrorNotReady = 600, hipEventDisableTiming = 2, hipStreamNonBlocking = 1 };

// Occupancy: eight resident workgroups per CU for every kernel (with two CUs,
// a partitioned launch over enough tiles gets sixteen workgroups, two per XCD).
template <typename F>
inline hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int* blocks, F, int, size_t)
{
    *blocks = 8;
    return hipSuccess;
}

#define hipLaunchKernelGGL(F, G, B, SH, ST, ...) emu::launch(G, B, [=]() { F(__VA_ARGS__); })
