// svdag_trace.hip -- primary rays through the flat layout.
//
// Counterpart of the reference's tracer kernel trace_paths
// (src/DAGTracer.cu:177-428): one ray per pixel from the pinhole camera of
// svdag_trace_params (src/DAGTracer.h:17-29), an ordered stack walk of the DAG
// that visits the children the ray crosses front to back, and per pixel the
// grid position of the first voxel hit plus the reference's debug colour.
//
// The arithmetic is the reference's, step for step: the ray direction in
// float64 (normalised, then rounded to float32), the node slab tests in
// float32.  Contraction into fused multiply-adds is off for this file, so the
// result is a function of IEEE operations alone and a host build of the same
// steps gives the same pixels.
//
// Layout per wave: 8x8 pixels (neighbouring rays walk the same nodes), four
// waves per workgroup (16x16 pixels).  The walk's stack (one entry per level:
// node word index, child mask, children still to visit) lives in LDS as a
// 32-bit index array and a 16-bit mask array (31.5 KB: five workgroups per
// CU), level-major so a wave's 64 entries of one level are consecutive.
#include "svdag_device.hpp"
#include "svdag_internal.hpp"

#pragma clang fp contract(off)

namespace svdag {
namespace {

constexpr int kTraceThreads = 256;
constexpr int kTraceTile = 16;  // pixels per workgroup side
constexpr int kTraceMaxDepth = 21;

// max and min with the reference's operand order and NaN behaviour
// ((a > b) ? a : b, src/CudaMath.h:12-15).
__device__ __forceinline__ float sel_max(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float sel_min(float a, float b) { return a < b ? a : b; }

struct Ray {
    float o[3], d[3], inv[3];
    uint32_t order;  // child-index bits to flip so children come front to back
};

// Children of the node at `level` with grid position `path` that the ray may
// cross (src/DAGTracer.cu:177-290): the child holding the midpoint of the
// ray's span inside the node, plus, for each of the node's three mid-planes the
// span crosses, the children on either side of the crossing point (both sides
// when the point is within 1e-4 of another mid-plane).  The root returns 0
// when the ray misses the volume.
__device__ __forceinline__ uint32_t crossed_children(bool root, int level, int depth, const uint32_t (&path)[3],
                                                     const Ray& r)
{
    const uint32_t shift = uint32_t(depth - level);
    const float radius = float(1u << (shift - 1));
    float rel[3], tm[3], lo[3], hi[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        rel[a] = (radius + float(path[a] << shift)) - r.o[a];
        tm[a] = rel[a] * r.inv[a];
        const float slab = radius * fabsf(r.inv[a]);
        lo[a] = tm[a] - slab;
        hi[a] = tm[a] + slab;
    }
    const float tmin = sel_max(sel_max(lo[0], sel_max(lo[1], lo[2])), 0.0f);
    const float tmax = sel_min(hi[0], sel_min(hi[1], hi[2]));
    if (root && tmin >= tmax) return 0;
    const float mid = 0.5f * (tmin + tmax);
    uint32_t m = 1u << ((mid * r.d[0] >= rel[0] ? 4u : 0u) + (mid * r.d[1] >= rel[1] ? 2u : 0u) +
                        (mid * r.d[2] >= rel[2] ? 1u : 0u));
    constexpr float eps = 1e-4f;
    // Per axis: the masks of the children above (hi) and below (lo) the
    // node's mid-plane of that axis.
    constexpr uint32_t kHi[3] = {0xF0u, 0xCCu, 0xAAu}, kLo[3] = {0x0Fu, 0x33u, 0x55u};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (!(tmin <= tm[a] && tm[a] <= tmax)) continue;
        const int b = a == 0 ? 1 : 0, c = a == 2 ? 1 : 2;  // the two other axes, in order
        const float pb = tm[a] * r.d[b], pc = tm[a] * r.d[c];
        uint32_t mb = 0, mc = 0;
        if (pb >= rel[b] - eps) mb |= kHi[b];
        if (pb <= rel[b] + eps) mb |= kLo[b];
        if (pc >= rel[c] - eps) mc |= kHi[c];
        if (pc <= rel[c] + eps) mc |= kLo[c];
        m |= mb & mc;
    }
    return m;
}

// First set child of `mask` in the ray's front-to-back order.
__device__ __forceinline__ uint32_t next_child(uint32_t order, uint32_t mask)
{
    // Children in visiting order are child ^ order for child = 0..7; the first
    // one set is the lowest set bit of the mask permuted by XOR with order.
    uint32_t permuted = 0;
#pragma unroll
    for (uint32_t c = 0; c < 8; ++c) permuted |= ((mask >> (c ^ order)) & 1u) << c;
    return uint32_t(__builtin_ctz(permuted)) ^ order;
}

__device__ __forceinline__ uint32_t murmur32(uint32_t h)
{
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}

__device__ __forceinline__ uint32_t rgb888(float r, float g, float b)
{
    const float R = fminf(1.0f, fmaxf(0.0f, r)), G = fminf(1.0f, fmaxf(0.0f, g)), B = fminf(1.0f, fmaxf(0.0f, b));
    return uint32_t(R * 255.0f) | uint32_t(G * 255.0f) << 8 | uint32_t(B * 255.0f) << 16;
}

__global__ __launch_bounds__(kTraceThreads) void k_trace_paths(const uint32_t* __restrict__ words, uint64_t n_words,
                                                               int depth, svdag_trace_params p,
                                                               uint32_t* __restrict__ paths,
                                                               uint32_t* __restrict__ colors)
{
    __shared__ uint32_t s_index[kTraceMaxDepth][kTraceThreads];
    __shared__ uint16_t s_masks[kTraceMaxDepth][kTraceThreads];  // child mask | children to visit << 8
    const uint32_t w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const uint32_t px = blockIdx.x * kTraceTile + (w & 1u) * 8u + (lane & 7u);
    const uint32_t py = blockIdx.y * kTraceTile + (w >> 1) * 8u + (lane >> 3);
    if (px >= p.width || py >= p.height) return;
    // Entry l of this thread's stack: stack_index[l * kTraceThreads], stack_masks[...].
    uint32_t* const stack_index = &s_index[0][threadIdx.x];
    uint16_t* const stack_masks = &s_masks[0][threadIdx.x];

    // Ray (src/DAGTracer.cu:308-319).
    Ray r;
    double dd[3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
        dd[a] = ((p.ray_min[a] + double(px) * p.ray_ddx[a]) + double(py) * p.ray_ddy[a]) - p.camera_position[a];
    const double inv_len = 1.0 / sqrt((dd[0] * dd[0] + dd[1] * dd[1]) + dd[2] * dd[2]);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        r.o[a] = float(p.camera_position[a]);
        r.d[a] = float(inv_len * dd[a]);
        r.inv[a] = float(1.0 / double(r.d[a]));
    }
    r.order = (r.d[0] < 0.0f ? 4u : 0u) + (r.d[1] < 0.0f ? 2u : 0u) + (r.d[2] < 0.0f ? 1u : 0u);

    const int leaf_level = depth - 2;
    uint32_t path[3] = {0, 0, 0};
    uint32_t debug_index = 0;
    bool bad = false;
    uint32_t index = 0, child_mask = 0, visit = 0;
    uint32_t leaf_lo = 0, leaf_hi = 0;
    if (n_words > 0) {
        child_mask = words[0] & 0xFFu;
        visit = child_mask & crossed_children(true, 0, depth, path, r);
    }
    int level = 0;
    // Each pass either clears one child of the current node or returns to an
    // ancestor, so the walk ends after finitely many passes.
    for (;;) {
        int up = level;
        while (up > 0 && !visit) {
            --up;
            index = stack_index[up * kTraceThreads];
            const uint32_t m = stack_masks[up * kTraceThreads];
            child_mask = m & 0xFFu;
            visit = m >> 8;
        }
        if (up == 0 && !visit) {
            path[0] = path[1] = path[2] = 0;
            break;
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) path[a] >>= uint32_t(level - up);
        level = up;

        const uint32_t child = next_child(r.order, visit);
        visit &= ~(1u << child);
        path[0] = path[0] << 1 | (child >> 2 & 1u);
        path[1] = path[1] << 1 | (child >> 1 & 1u);
        path[2] = path[2] << 1 | (child & 1u);
        stack_index[level * kTraceThreads] = index;
        stack_masks[level * kTraceThreads] = uint16_t(child_mask | visit << 8);
        ++level;
        if (uint32_t(depth - 2 - level) == p.debug_level) debug_index = index;
        if (level == depth) break;  // a voxel

        if (level <= leaf_level) {
            // Interior node: its word, then one word index per set child.
            const uint64_t at = uint64_t(index) + 1 + uint32_t(__popc(child_mask & ((1u << child) - 1u)));
            if (at >= n_words) {
                bad = true;
                break;
            }
            const uint32_t next = words[at];
            if (level < leaf_level) {
                if (next >= n_words) {
                    bad = true;
                    break;
                }
                index = next;
                child_mask = words[next] & 0xFFu;
            } else {
                // A leaf: two words of 4x4x4 occupancy; byte i of the 64 bits
                // is child i's 2x2x2 block.
                if (uint64_t(next) + 1 >= n_words) {
                    bad = true;
                    break;
                }
                leaf_lo = words[next];
                leaf_hi = words[next + 1];
                child_mask = 0;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    child_mask |= ((leaf_lo >> (8 * i)) & 0xFFu) ? 1u << i : 0u;
                    child_mask |= ((leaf_hi >> (8 * i)) & 0xFFu) ? 1u << (i + 4) : 0u;
                }
            }
        } else {
            // Inside a leaf: the child's byte of the occupancy.
            child_mask = ((child & 4u) ? leaf_hi : leaf_lo) >> ((child & 3u) * 8u) & 0xFFu;
        }
        visit = child_mask & crossed_children(false, level, depth, path, r);
    }

    const uint64_t pix = uint64_t(py) * p.width + px;
    if (bad) path[0] = path[1] = path[2] = 0xFFFFFFFFu;
    if (paths) {
        paths[3 * pix + 0] = path[0];
        paths[3 * pix + 1] = path[1];
        paths[3 * pix + 2] = path[2];
    }
    if (colors) {
        uint32_t c;
        if (bad) c = 0;
        else if (p.debug_level == 0xFFFFFFFFu)
            c = rgb888(float(path[0] & 0xFFu) / 255.0f, float(path[1] & 0xFFu) / 255.0f,
                       float(path[2] & 0xFFu) / 255.0f);
        else c = murmur32(debug_index);
        colors[pix] = c;
    }
}

}  // namespace

static_assert(kTraceMaxDepth >= 21, "stack holds every level of the deepest DAG");

void launch_trace_paths(hipStream_t s, const uint32_t* words, uint64_t n_words, int depth,
                        const svdag_trace_params& p, uint32_t* paths, uint32_t* colors)
{
    if (p.width == 0 || p.height == 0) return;
    TimedLaunch t(kKTracePaths, s);
    const dim3 grid((p.width + kTraceTile - 1) / kTraceTile, (p.height + kTraceTile - 1) / kTraceTile);
    hipLaunchKernelGGL(k_trace_paths, grid, dim3(kTraceThreads), 0, s, words, n_words, depth, p, paths, colors);
}

}  // namespace svdag
