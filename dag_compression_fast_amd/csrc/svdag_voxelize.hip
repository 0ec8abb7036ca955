// svdag_voxelize.hip -- device fragment producer: a triangle soup -> unsorted
// Morton keys with duplicates, the input contract of the DAG build.
//
// The reference produces its fragments with an OpenGL conservative rasteriser
// (src/Voxelizer.cpp:254-369, Shaders/FragmentShader.glsl:19-32, 208-209).
// Here the triangles are voxelised by the thin surface rule of the
// triangle-surface workload generator (after Schwarz & Seidel 2010): one voxel
// per column along the dominant normal axis, kept when the voxel centre passes
// the three projected edge tests dilated by half the larger edge-normal
// component.  Every floating-point step is the generator's, in the same order
// and without fused multiply-adds, so both emit the same voxels.
//
// Work is spread by columns, not triangles (sizes are heavy-tailed):
//   k_tri_setup     per triangle: normal, dominant axis, edge functions and
//                   the column rectangle; its column count
//   (scan)          column offsets per triangle
//   k_vox_count     tile of 2048 global columns: voxels emitted
//   (scan)          key offsets per tile
//   k_vox_write     the same columns again, writing the keys
#pragma clang fp contract(off)

#include "svdag_device.hpp"
#include "svdag_internal.hpp"

namespace svdag {

namespace {

constexpr int kVoxThreads = 256;
constexpr int kVoxIpt = 8;
constexpr uint64_t kVoxTile = uint64_t(kVoxThreads) * kVoxIpt;
static_assert(kVoxTile == kVoxTileItems, "tile size shared with the planner");

__device__ __forceinline__ double pick3(const double (&v)[3], int i) { return i == 0 ? v[0] : (i == 1 ? v[1] : v[2]); }

__device__ __forceinline__ uint64_t split_by_3(uint32_t a)
{
    uint64_t x = a & 0x1fffff;
    x = (x | x << 32) & 0x1f00000000ffffull;
    x = (x | x << 16) & 0x1f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}

// Morton layout of src/DAG_Compression_Fast.cpp:223-236.
__device__ __forceinline__ uint64_t morton3(uint32_t x, uint32_t y, uint32_t z)
{
    return split_by_3(x) << 2 | split_by_3(y) << 1 | split_by_3(z);
}

__global__ __launch_bounds__(kVoxThreads) void k_tri_setup(const double* __restrict__ tris, uint64_t n_tris,
                                                           int64_t side, TriSetup* __restrict__ setup,
                                                           uint32_t* __restrict__ cols,
                                                           unsigned long long* __restrict__ err)
{
    const uint64_t t = uint64_t(blockIdx.x) * kVoxThreads + threadIdx.x;
    if (t >= n_tris) return;
    double vs[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int k = 0; k < 3; ++k) vs[c][k] = tris[9 * t + 3 * c + k];
    // n = (v1 - v0) x (v2 - v0)
    const double e1x = vs[1][0] - vs[0][0], e1y = vs[1][1] - vs[0][1], e1z = vs[1][2] - vs[0][2];
    const double e2x = vs[2][0] - vs[0][0], e2y = vs[2][1] - vs[0][1], e2z = vs[2][2] - vs[0][2];
    TriSetup S{};
    S.n[0] = e1y * e2z - e1z * e2y;
    S.n[1] = e1z * e2x - e1x * e2z;
    S.n[2] = e1x * e2y - e1y * e2x;
    const double an0 = fabs(S.n[0]), an1 = fabs(S.n[1]), an2 = fabs(S.n[2]);
    const int a = an0 >= an1 && an0 >= an2 ? 0 : (an1 >= an2 ? 1 : 2);
    const double ana = a == 0 ? an0 : (a == 1 ? an1 : an2);
    if (ana == 0.0) {  // degenerate
        setup[t] = S;
        cols[t] = 0;
        return;
    }
    const int u = (a + 1) % 3, w = (a + 2) % 3;
    S.a = a;
    S.u = u;
    S.w = w;
    // Projected edge functions for the three axis planes (xy, yz, zx).
    for (int p = 0; p < 3; ++p) {
        const int i0 = p, i1 = (p + 1) % 3;
        const double sgn = pick3(S.n, (p + 2) % 3) >= 0 ? 1.0 : -1.0;
        if (pick3(S.n, (p + 2) % 3) != 0.0) S.planes |= 1 << p;
        for (int e = 0; e < 3; ++e) {
            const int b = (e + 1) % 3;
            const double ex = vs[b][i0] - vs[e][i0], ey = vs[b][i1] - vs[e][i1];
            const double nx = -ey * sgn, ny = ex * sgn;
            S.ex[p][e] = nx;
            S.ey[p][e] = ny;
            S.ed[p][e] = -(nx * vs[e][i0] + ny * vs[e][i1]) + 0.5 * fmax(fabs(nx), fabs(ny));
        }
    }
    const double lo_u = fmin(fmin(vs[0][u], vs[1][u]), vs[2][u]), hi_u = fmax(fmax(vs[0][u], vs[1][u]), vs[2][u]);
    const double lo_w = fmin(fmin(vs[0][w], vs[1][w]), vs[2][w]), hi_w = fmax(fmax(vs[0][w], vs[1][w]), vs[2][w]);
    const int64_t u0 = max(int64_t(0), int64_t(floor(lo_u)) - 1), u1 = min(side - 1, int64_t(floor(hi_u)) + 1);
    const int64_t w0 = max(int64_t(0), int64_t(floor(lo_w)) - 1), w1 = min(side - 1, int64_t(floor(hi_w)) + 1);
    S.nd = S.n[0] * vs[0][0] + S.n[1] * vs[0][1] + S.n[2] * vs[0][2];
    S.u0 = u0;
    S.w0 = w0;
    S.nw = w1 >= w0 ? uint32_t(w1 - w0 + 1) : 0u;
    const uint64_t nu = u1 >= u0 ? uint64_t(u1 - u0 + 1) : 0u;
    const uint64_t c = nu * S.nw;
    setup[t] = S;
    cols[t] = c <= 0xFFFFFFFFull ? uint32_t(c) : 0u;
    if (c > 0xFFFFFFFFull) atomicOr(err, (unsigned long long)kErrInternal);  // more columns than a u32 count holds
}

// Voxels of column (iu, iw) of triangle S (at most two), appended to k[] at
// `found`; returns the new count.
template <int N>
__device__ __forceinline__ int column_voxels(const TriSetup& S, int64_t iu, int64_t iw, int64_t side, uint64_t (&k)[N],
                                             int found)
{
    const double cu = double(iu) + 0.5, cw = double(iw) + 0.5;
    const double na = pick3(S.n, S.a), nu = pick3(S.n, S.u), nw = pick3(S.n, S.w);
    const double ac = (S.nd - nu * cu - nw * cw) / na;
    const int64_t k0 = int64_t(floor(ac - 1.0)), k1 = int64_t(floor(ac));
    for (int64_t kk = max(k0, int64_t(0)); kk <= min(k1, side - 1); ++kk) {
        const double ck = double(kk) + 0.5;
        if (fabs(ck - ac) > 0.5) continue;
        // Voxel centre: k + 1/2 along a, cu along u, cw along w.
        const double cc[3] = {S.a == 0 ? ck : (S.u == 0 ? cu : cw), S.a == 1 ? ck : (S.u == 1 ? cu : cw),
                              S.a == 2 ? ck : (S.u == 2 ? cu : cw)};
        bool inside = true;
        for (int p = 0; p < 3 && inside; ++p) {
            if (!((S.planes >> p) & 1)) continue;  // triangle is edge-on in this projection
            const double c0 = pick3(cc, p), c1 = pick3(cc, (p + 1) % 3);
            for (int e = 0; e < 3; ++e)
                if (S.ex[p][e] * c0 + S.ey[p][e] * c1 + S.ed[p][e] < 0) {
                    inside = false;
                    break;
                }
        }
        if (!inside) continue;
        const int64_t x = S.a == 0 ? kk : (S.u == 0 ? iu : iw);
        const int64_t y = S.a == 1 ? kk : (S.u == 1 ? iu : iw);
        const int64_t z = S.a == 2 ? kk : (S.u == 2 ? iu : iw);
        k[found++] = morton3(uint32_t(x), uint32_t(y), uint32_t(z));
    }
    return found;
}

// Triangle holding global column g: the last t in [lo, hi) with off[t] <= g.
__device__ __forceinline__ uint64_t triangle_of(const uint64_t* __restrict__ off, uint64_t lo, uint64_t hi, uint64_t g)
{
    while (hi - lo > 1) {
        const uint64_t mid = (lo + hi) / 2;
        if (off[mid] <= g) lo = mid;
        else hi = mid;
    }
    return lo;
}

// The triangles of tile t's columns lie in [tile_tri[t], tile_tri[t + 1] + 1)
// (tile_tri[t] = triangle of the tile's first column, one entry past the last
// tile), found by one thread per tile.
__global__ __launch_bounds__(kVoxThreads) void k_tile_tris(const uint64_t* __restrict__ off,
                                                           const uint64_t* __restrict__ n_tris_ptr,
                                                           const uint64_t* __restrict__ total_ptr,
                                                           uint64_t* __restrict__ tile_tri)
{
    const uint64_t n_tris = *n_tris_ptr, total = *total_ptr;
    const uint64_t ntiles = (total + kVoxTile - 1) / kVoxTile;
    for (uint64_t t = uint64_t(blockIdx.x) * kVoxThreads + threadIdx.x; t <= ntiles;
         t += uint64_t(gridDim.x) * kVoxThreads)
        tile_tri[t] = t < ntiles ? triangle_of(off, 0, n_tris, t * kVoxTile) : n_tris - 1;
}

// Keys of the thread's kVoxIpt consecutive columns of tile `tile` (at most two
// per column) into k[]; returns how many.  One triangle search and one
// division per thread: the column position then steps along w, then u.
constexpr int kVoxKeys = 2 * kVoxIpt;

__device__ __forceinline__ int thread_columns(const TriSetup* __restrict__ setup, const uint64_t* __restrict__ off,
                                              const uint64_t* __restrict__ tile_tri, uint64_t n_tris, uint64_t total,
                                              uint64_t tile, int64_t side, uint64_t (&k)[kVoxKeys])
{
    const uint64_t g0 = tile * kVoxTile + uint64_t(threadIdx.x) * kVoxIpt;
    if (g0 >= total) return 0;
    const uint64_t g1 = min(g0 + uint64_t(kVoxIpt), total);
    const uint64_t hi = tile_tri[tile + 1] + 1;
    uint64_t t = triangle_of(off, tile_tri[tile], hi, g0);
    uint64_t t_end = t + 1 < n_tris ? off[t + 1] : total;
    uint32_t nw = setup[t].nw;
    const uint32_t j = uint32_t(g0 - off[t]);  // a triangle has fewer than 2^32 columns
    uint32_t ju = j / nw, jw = j - ju * nw;
    int found = 0;
#pragma unroll 1
    for (uint64_t g = g0; g < g1; ++g) {
        if (g >= t_end) {  // the next triangle with columns starts at g
            t = triangle_of(off, t + 1, hi, g);
            t_end = t + 1 < n_tris ? off[t + 1] : total;
            nw = setup[t].nw;
            ju = 0;
            jw = 0;
        }
        const TriSetup& S = setup[t];
        found = column_voxels(S, S.u0 + int64_t(ju), S.w0 + int64_t(jw), side, k, found);
        if (++jw == nw) {
            jw = 0;
            ++ju;
        }
    }
    return found;
}

__global__ __launch_bounds__(kVoxThreads) void k_vox_count(const TriSetup* __restrict__ setup,
                                                           const uint64_t* __restrict__ off,
                                                           const uint64_t* __restrict__ tile_tri,
                                                           const uint64_t* __restrict__ n_tris_ptr,
                                                           const uint64_t* __restrict__ total_ptr, int64_t side,
                                                           uint32_t* __restrict__ counts)
{
    __shared__ uint64_t s_red[kVoxThreads / kWave];
    const uint64_t n_tris = *n_tris_ptr, total = *total_ptr;
    const uint64_t ntiles = (total + kVoxTile - 1) / kVoxTile;
    for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        uint64_t k[kVoxKeys];
        const int mine = thread_columns(setup, off, tile_tri, n_tris, total, tile, side, k);
        const uint64_t sum = block_reduce_sum<kVoxThreads>(uint64_t(mine), s_red);
        if (threadIdx.x == 0) counts[tile] = uint32_t(sum);
    }
}

__global__ __launch_bounds__(kVoxThreads) void k_vox_write(const TriSetup* __restrict__ setup,
                                                           const uint64_t* __restrict__ off,
                                                           const uint64_t* __restrict__ tile_tri,
                                                           const uint64_t* __restrict__ n_tris_ptr,
                                                           const uint64_t* __restrict__ total_ptr, int64_t side,
                                                           const uint64_t* __restrict__ tile_offsets,
                                                           uint64_t* __restrict__ keys)
{
    __shared__ uint32_t s_scan[kVoxThreads / kWave];
    const uint64_t n_tris = *n_tris_ptr, total = *total_ptr;
    const uint64_t ntiles = (total + kVoxTile - 1) / kVoxTile;
    for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        uint64_t k[kVoxKeys];
        const int mine = thread_columns(setup, off, tile_tri, n_tris, total, tile, side, k);
        uint32_t unused;
        const uint64_t o = tile_offsets[tile] + block_exclusive_scan<kVoxThreads>(uint32_t(mine), s_scan, unused);
#pragma unroll
        for (int j = 0; j < kVoxKeys; ++j)
            if (j < mine) keys[o + j] = k[j];
    }
}

}  // namespace

void launch_tri_setup(hipStream_t s, const double* tris, uint64_t n_tris, int depth, TriSetup* setup, uint32_t* cols,
                      unsigned long long* err)
{
    if (n_tris == 0) return;
    TimedLaunch t(kKVoxSetup, s);
    hipLaunchKernelGGL(k_tri_setup, dim3(uint32_t((n_tris + kVoxThreads - 1) / kVoxThreads)), dim3(kVoxThreads), 0, s,
                       tris, n_tris, int64_t(1) << depth, setup, cols, err);
}

void launch_vox_count(hipStream_t s, const TriSetup* setup, const uint64_t* col_offsets, const uint64_t* n_tris_dev,
                      const uint64_t* total_cols_dev, uint64_t total_cols_bound, int depth, uint64_t* tile_tri,
                      uint32_t* tile_counts)
{
    TimedLaunch t(kKVoxCount, s);
    const uint64_t tiles = (total_cols_bound + kVoxTile - 1) / kVoxTile;
    const int grid = int(std::max<uint64_t>(1, std::min<uint64_t>(tiles, uint64_t(device_cu_count()) * 8)));
    hipLaunchKernelGGL(k_tile_tris, dim3(uint32_t((tiles + kVoxThreads) / kVoxThreads)), dim3(kVoxThreads), 0, s,
                       col_offsets, n_tris_dev, total_cols_dev, tile_tri);
    hipLaunchKernelGGL(k_vox_count, dim3(grid), dim3(kVoxThreads), 0, s, setup, col_offsets, tile_tri, n_tris_dev,
                       total_cols_dev, int64_t(1) << depth, tile_counts);
}

void launch_vox_write(hipStream_t s, const TriSetup* setup, const uint64_t* col_offsets, const uint64_t* tile_tri,
                      const uint64_t* n_tris_dev, const uint64_t* total_cols_dev, uint64_t total_cols_bound, int depth,
                      const uint64_t* tile_offsets, uint64_t* keys)
{
    TimedLaunch t(kKVoxWrite, s);
    const uint64_t tiles = (total_cols_bound + kVoxTile - 1) / kVoxTile;
    const int grid = int(std::max<uint64_t>(1, std::min<uint64_t>(tiles, uint64_t(device_cu_count()) * 8)));
    hipLaunchKernelGGL(k_vox_write, dim3(grid), dim3(kVoxThreads), 0, s, setup, col_offsets, tile_tri, n_tris_dev,
                       total_cols_dev, int64_t(1) << depth, tile_offsets, keys);
}

}  // namespace svdag
