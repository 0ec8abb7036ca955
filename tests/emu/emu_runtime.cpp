// CPU kernel emulator runtime (tests only): fiber scheduler for the mock
// <hip/hip_runtime.h> and host-side stand-ins for the HIP runtime calls.
#include "hip/hip_runtime.h"

#include <signal.h>
#if defined(__SANITIZE_ADDRESS__)
#include <sanitizer/asan_interface.h>
#endif

#include <atomic>
#include <cstdint>
#include <cstring>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <thread>
#include <unistd.h>
#include <vector>

// Stack switch between fibers (x86-64 System V): saves the callee-saved
// registers on the current stack, stores its stack pointer in *from and resumes
// the stack `to`.  A user-space switch: no signal-mask system call per switch.
extern "C" void emu_switch(void** from, void* to);
asm(R"(
    .text
    .globl emu_switch
    .type emu_switch, @function
emu_switch:
    pushq %rbp
    pushq %rbx
    pushq %r12
    pushq %r13
    pushq %r14
    pushq %r15
    movq %rsp, (%rdi)
    movq %rsi, %rsp
    popq %r15
    popq %r14
    popq %r13
    popq %r12
    popq %rbx
    popq %rbp
    ret
    .size emu_switch, .-emu_switch
)");

namespace emu {
namespace {

struct Fiber {
    void* sp = nullptr;
    dim3 tid;
    bool done = false;
};

struct Wave {
    unsigned count = 0, gen = 0, size = 64;
    unsigned long long slot[64];
};

constexpr size_t kStack = 512 * 1024;

// Per OS thread (= per running block):
thread_local void* g_main_sp = nullptr;
thread_local std::vector<Fiber> g_fibers;
// Fiber stacks: each launch starts fresh worker threads; a thread takes stacks
// from the process-wide free list and returns them when it ends.
std::mutex g_free_mutex;
std::vector<char*> g_free_stacks;

struct StackPool {
    std::vector<char*> stacks;
    ~StackPool()
    {
        std::lock_guard<std::mutex> lock(g_free_mutex);
        g_free_stacks.insert(g_free_stacks.end(), stacks.begin(), stacks.end());
    }
};

char* take_stack()
{
    {
        std::lock_guard<std::mutex> lock(g_free_mutex);
        if (!g_free_stacks.empty()) {
            char* p = g_free_stacks.back();
            g_free_stacks.pop_back();
            return p;
        }
    }
    return static_cast<char*>(std::malloc(kStack));
}
thread_local StackPool g_stack_pool;
thread_local std::vector<char*>& g_stacks = g_stack_pool.stacks;
thread_local std::vector<Wave> g_waves;
thread_local int g_cur = -1;
thread_local dim3 g_block;
thread_local unsigned g_bar_count = 0, g_bar_gen = 0;
thread_local std::chrono::steady_clock::time_point g_block_start;
// Per launch:
dim3 g_block_dim, g_grid_dim;
std::function<void()>* g_body = nullptr;
int g_launch_threads = 8;

void yield()
{
    const int me = g_cur;
    emu_switch(&g_fibers[me].sp, g_main_sp);
}

// Every lane, the last arriver included, yields once the wave has met, so the
// lanes resume in lane order like one lock-stepped instruction stream.
void wave_barrier(Wave& w)
{
    const unsigned gen = w.gen;
    if (++w.count == w.size) {
        w.count = 0;
        ++w.gen;
        yield();
    } else {
        while (w.gen == gen) yield();
    }
}

unsigned flat_tid() { return g_fibers[g_cur].tid.x + g_block_dim.x * (g_fibers[g_cur].tid.y + g_block_dim.y * g_fibers[g_cur].tid.z); }

Wave& my_wave() { return g_waves[flat_tid() / 64]; }

[[noreturn]] void fiber_entry()
{
    (*g_body)();
    g_fibers[g_cur].done = true;
    void* dead = nullptr;
    emu_switch(&dead, g_main_sp);
    std::abort();  // a finished fiber is never resumed
}

// A fresh stack whose first switch-in "returns" into fiber_entry with the
// stack aligned as after a call.
void* fresh_stack(char* base)
{
#if defined(__SANITIZE_ADDRESS__)
    // A fiber that ended never returned from its frames: under AddressSanitizer
    // their stack redzones are still poisoned when the stack is reused.
    __asan_unpoison_memory_region(base, kStack);
#endif
    uintptr_t top = (reinterpret_cast<uintptr_t>(base) + kStack) & ~uintptr_t(15);
    void** sp = reinterpret_cast<void**>(top - 64);
    for (int i = 0; i < 6; ++i) sp[i] = nullptr;  // r15 r14 r13 r12 rbx rbp
    sp[6] = reinterpret_cast<void*>(&fiber_entry);
    return sp;
}

// Names the block, the lane and the faulting address when emulated kernel code
// faults (runs on an alternate stack, so a fiber stack overflow reports too).
void on_fault(int sig, siginfo_t* info, void*)
{
    char buf[256];
    const unsigned t = g_cur >= 0 ? g_fibers[g_cur].tid.x : ~0u;
    const char* base = g_cur >= 0 ? g_stacks[g_cur] : nullptr;
    const int len = std::snprintf(buf, sizeof(buf),
                                  "emu: signal %d at address %p in block %u thread %u (fiber stack %p..%p)\n", sig,
                                  info->si_addr, g_block.x, t, static_cast<const void*>(base),
                                  static_cast<const void*>(base ? base + kStack : nullptr));
    if (len > 0) (void)!write(2, buf, size_t(len));
    signal(sig, SIG_DFL);
    raise(sig);
}

void install_fault_reporter()
{
#if defined(__SANITIZE_ADDRESS__)
    return;  // AddressSanitizer reports faults itself and owns the signal stack
#endif
    thread_local bool done = false;
    if (done) return;
    done = true;
    stack_t ss{};
    ss.ss_size = 1 << 16;
    ss.ss_sp = std::malloc(ss.ss_size);
    sigaltstack(&ss, nullptr);
    struct sigaction sa{};
    sa.sa_sigaction = on_fault;
    sa.sa_flags = SA_SIGINFO | SA_ONSTACK;
    sigaction(SIGSEGV, &sa, nullptr);
    sigaction(SIGBUS, &sa, nullptr);
}

void run_block(unsigned nthreads)
{
    install_fault_reporter();
    if (g_stacks.size() < nthreads)
        while (g_stacks.size() < nthreads) {
            char* p = take_stack();
            if (!p) {
                std::fprintf(stderr, "emu: cannot allocate a fiber stack\n");
                std::abort();
            }
            g_stacks.push_back(p);
        }
    g_fibers.assign(nthreads, Fiber{});
    g_waves.assign((nthreads + 63) / 64, Wave{});
    for (size_t w = 0; w < g_waves.size(); ++w) g_waves[w].size = std::min<unsigned>(64, nthreads - unsigned(w) * 64);
    g_bar_count = 0;
    for (unsigned i = 0; i < nthreads; ++i) {
        Fiber& f = g_fibers[i];
        f.tid = dim3(i % g_block_dim.x, (i / g_block_dim.x) % g_block_dim.y, i / (g_block_dim.x * g_block_dim.y));
        f.sp = fresh_stack(g_stacks[i]);
    }
    unsigned alive = nthreads;
    while (alive) {
        for (unsigned i = 0; i < nthreads; ++i) {
            if (g_fibers[i].done) continue;
            g_cur = int(i);
            emu_switch(&g_main_sp, g_fibers[i].sp);
            if (g_fibers[i].done) --alive;
        }
    }
    g_cur = -1;
}

}  // namespace

const dim3& thread_idx() { return g_fibers[g_cur].tid; }
const dim3& block_idx() { return g_block; }
const dim3& block_dim() { return g_block_dim; }
const dim3& grid_dim() { return g_grid_dim; }
unsigned lane() { return flat_tid() % 64; }

void sync_block()
{
    const unsigned gen = g_bar_gen;
    if (++g_bar_count == g_fibers.size()) {
        g_bar_count = 0;
        ++g_bar_gen;
    } else {
        while (g_bar_gen == gen) yield();
    }
}

void yield_sleep()
{
    if (std::chrono::steady_clock::now() - g_block_start > std::chrono::seconds(60)) {
        std::fprintf(stderr, "emu: a spin-wait did not resolve in 60 s (block %u thread %u)\n", g_block.x, flat_tid());
        std::abort();
    }
    yield();
}

void wave_sync() { wave_barrier(my_wave()); }

unsigned long long ballot(bool p)
{
    Wave& w = my_wave();
    w.slot[lane()] = p ? 1 : 0;
    wave_barrier(w);
    unsigned long long r = 0;
    for (unsigned i = 0; i < w.size; ++i)
        if (w.slot[i]) r |= 1ull << i;
    wave_barrier(w);
    return r;
}

unsigned long long exchange(unsigned long long v, int src)
{
    Wave& w = my_wave();
    w.slot[lane()] = v;
    wave_barrier(w);
    const unsigned long long r = (src >= 0 && unsigned(src) < w.size) ? w.slot[src] : v;
    wave_barrier(w);
    return r;
}

std::mutex g_launch_mutex;  // launches from several host threads run one at a time

// Block schedule: 0 = blocks taken in increasing order by a pool of OS threads
// (concurrent); 1 = one block at a time in decreasing order.  Under the second,
// a block that waited for a lower-numbered block of the same launch would never
// see it run (the spin guard aborts), and a block that read what another block
// of the launch writes would read it before it is written.
int g_schedule = 0;
// Launch-configuration errors, reported by the next hipGetLastError (as the HIP
// runtime reports an invalid configuration) instead of running the kernel.
int g_sticky_error = 0;
uint64_t g_launches = 0, g_bad_launches = 0;

bool valid_config(dim3 grid, dim3 block)
{
    const uint64_t threads = uint64_t(block.x) * block.y * block.z;
    return grid.x >= 1 && grid.y >= 1 && grid.z >= 1 && grid.x <= 0x7FFFFFFFu && grid.y <= 65535u &&
           grid.z <= 65535u && threads >= 1 && threads <= 1024 && block.z <= 1024;
}

// Stream capture: operations recorded instead of run (one capture at a time).
struct GraphOp {
    dim3 grid, block;
    std::function<void()> body;  // a kernel launch, or
    void* dst = nullptr;         // a copy
    const void* src = nullptr;
    size_t bytes = 0;
};
struct Graph {
    std::vector<GraphOp> ops;
};
Graph* g_capture = nullptr;

void run_launch(dim3 grid, dim3 block, std::function<void()>& body);

void launch(dim3 grid, dim3 block, std::function<void()> body)
{
    if (g_capture) {
        GraphOp op;
        op.grid = grid;
        op.block = block;
        op.body = std::move(body);
        g_capture->ops.push_back(std::move(op));
        return;
    }
    run_launch(grid, block, body);
}

void run_launch(dim3 grid, dim3 block, std::function<void()>& body)
{
    std::lock_guard<std::mutex> serial(g_launch_mutex);
    if (g_cur >= 0) {
        std::fprintf(stderr, "emu: launch from inside a kernel\n");
        std::abort();
    }
    ++g_launches;
    if (!valid_config(grid, block)) {
        std::fprintf(stderr, "emu: invalid launch configuration grid (%u, %u, %u) block (%u, %u, %u)\n", grid.x,
                     grid.y, grid.z, block.x, block.y, block.z);
        ++g_bad_launches;
        g_sticky_error = 9;  // hipErrorInvalidConfiguration
        return;
    }
    g_grid_dim = grid;
    g_block_dim = block;
    g_body = &body;
    const unsigned nthreads = block.x * block.y * block.z;
    const unsigned nblocks = grid.x * grid.y * grid.z;
    std::atomic<unsigned> next{0};
    const bool reverse = g_schedule == 1;
    auto worker = [&]() {
        for (;;) {
            const unsigned i = next.fetch_add(1);
            if (i >= nblocks) break;
            const unsigned b = reverse ? nblocks - 1 - i : i;
            g_block = dim3(b % grid.x, (b / grid.x) % grid.y, b / (grid.x * grid.y));
            g_block_start = std::chrono::steady_clock::now();
            run_block(nthreads);
        }
    };
    const unsigned workers = reverse ? 1u : std::min<unsigned>(nblocks, unsigned(g_launch_threads));
    std::vector<std::thread> pool;
    for (unsigned t = 1; t < workers; ++t) pool.emplace_back(worker);
    worker();
    for (auto& t : pool) t.join();
    g_body = nullptr;
}

}  // namespace emu

extern "C" {

static void* emu_alloc(size_t n)
{
    const size_t bytes = (n + 255) / 256 * 256;
    void* p = std::aligned_alloc(256, bytes ? bytes : 256);
    if (p) std::memset(p, 0xCD, bytes ? bytes : 256);  // device memory is never assumed zero
    return p;
}

hipError_t hipMalloc(void** p, size_t n) { *p = emu_alloc(n); return *p ? hipSuccess : hipErrorOutOfMemory; }
hipError_t hipFree(void* p) { std::free(p); return hipSuccess; }
// Device globals are host globals here: the symbol is its own address.
hipError_t hipGetSymbolAddress(void** p, const void* symbol)
{
    *p = const_cast<void*>(symbol);
    return hipSuccess;
}
hipError_t hipMallocAsync(void** p, size_t n, hipStream_t) { return hipMalloc(p, n); }
hipError_t hipFreeAsync(void* p, hipStream_t) { return hipFree(p); }
hipError_t hipHostMalloc(void** p, size_t n, unsigned) { return hipMalloc(p, n); }
hipError_t hipHostFree(void* p) { return hipFree(p); }
hipError_t hipMemcpy(void* d, const void* s, size_t n, int) { if (n) std::memmove(d, s, n); return hipSuccess; }
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, int k, hipStream_t)
{
    if (emu::g_capture) {
        emu::GraphOp op;
        op.dst = d;
        op.src = s;
        op.bytes = n;
        emu::g_capture->ops.push_back(std::move(op));
        return hipSuccess;
    }
    return hipMemcpy(d, s, n, k);
}
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode)
{
    if (emu::g_capture) return hipErrorInvalidValue;
    emu::g_capture = new emu::Graph();
    return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g)
{
    if (!emu::g_capture) return hipErrorInvalidValue;
    *g = emu::g_capture;
    emu::g_capture = nullptr;
    return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, hipGraphNode_t*, char*, size_t)
{
    *e = new emu::Graph(*static_cast<emu::Graph*>(g));
    return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t)
{
    for (emu::GraphOp& op : static_cast<emu::Graph*>(e)->ops) {
        if (op.body) emu::run_launch(op.grid, op.block, op.body);
        else if (op.bytes) std::memmove(op.dst, op.src, op.bytes);
    }
    return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t e)
{
    delete static_cast<emu::Graph*>(e);
    return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g)
{
    delete static_cast<emu::Graph*>(g);
    return hipSuccess;
}
hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t) { std::memset(d, v, n); return hipSuccess; }
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipDeviceSynchronize() { return hipSuccess; }
hipError_t hipGetLastError()
{
    const hipError_t e = emu::g_sticky_error;
    emu::g_sticky_error = 0;
    return e;
}
hipError_t hipPeekAtLastError() { return emu::g_sticky_error; }
void emu_set_schedule(int mode) { emu::g_schedule = mode; }
uint64_t emu_launch_count() { return emu::g_launches; }
uint64_t emu_bad_launch_count() { return emu::g_bad_launches; }
const char* hipGetErrorString(hipError_t) { return "emulator error"; }
hipError_t hipGetDevice(int* d) { *d = 0; return hipSuccess; }
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipGetDeviceCount(int* n) { *n = 1; return hipSuccess; }
hipError_t hipGetDeviceProperties(hipDeviceProp_t* p, int)
{
    p->multiProcessorCount = 2;
    std::strcpy(p->name, "cpu-emulator");
    return hipSuccess;
}
hipError_t hipDeviceGetDefaultMemPool(hipMemPool_t* p, int) { *p = nullptr; return hipSuccess; }
hipError_t hipMemPoolSetAttribute(hipMemPool_t, int, void*) { return hipSuccess; }
hipError_t hipMemPoolTrimTo(hipMemPool_t, size_t) { return hipSuccess; }
hipError_t hipEventCreate(void** e) { *e = std::malloc(8); return hipSuccess; }
hipError_t hipEventCreateWithFlags(void** e, unsigned) { *e = std::malloc(8); return hipSuccess; }
hipError_t hipEventRecord(void*, void*) { return hipSuccess; }
hipError_t hipEventQuery(void*) { return hipSuccess; }
hipError_t hipStreamQuery(void*) { return hipSuccess; }
hipError_t hipEventSynchronize(void*) { return hipSuccess; }
hipError_t hipEventElapsedTime(float* ms, void*, void*) { *ms = 0; return hipSuccess; }
hipError_t hipEventDestroy(void* e) { std::free(e); return hipSuccess; }
hipError_t hipStreamCreate(void** s) { *s = nullptr; return hipSuccess; }
// Launches run synchronously in issue order, so streams need no state; a
// created stream only needs a distinct non-null handle.
hipError_t hipStreamCreateWithFlags(void** s, unsigned) { *s = std::malloc(8); return hipSuccess; }
hipError_t hipStreamWaitEvent(void*, void*, unsigned) { return hipSuccess; }
hipError_t hipStreamDestroy(void*) { return hipSuccess; }
}
