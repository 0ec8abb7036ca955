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
//                   the column rectangle; its row count (rows run along w)
//   (scan)          row offsets per triangle
//   k_row_setup     per row: the columns whose centres can pass the dominant
//                   projection's edge tests (a conservative superset, so the
//                   exact per-column test below decides; about a fifth of the
//                   rectangle for the triangle workload)
//   (scan)          column offsets per row
//   k_vox_emit      tile of 2048 global columns: its keys (at most two per
//                   column) into the tile's 4096-key staging region, and
//                   their count
//   (scan)          key offsets per tile
//   k_vox_compact   each tile's staged keys to its offset
// When the staging area would be too large, k_vox_count counts the keys per
// tile and k_vox_write evaluates the columns again to write them in place.
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

// A triangle with a non-finite coordinate gets no rows and sets kErrKeyRange
// in *err.
__global__ __launch_bounds__(kVoxThreads) void k_tri_setup(const double* __restrict__ tris, uint64_t n_tris,
                                                           int64_t side, TriSetup* __restrict__ setup,
                                                           uint32_t* __restrict__ rows,
                                                           unsigned long long* __restrict__ rect_columns,
                                                           unsigned long long* __restrict__ err)
{
    __shared__ uint64_t s_red[kVoxThreads / kWave];
    const uint64_t t = uint64_t(blockIdx.x) * kVoxThreads + threadIdx.x;
    uint64_t rect = 0;  // columns of the triangle's rectangle (bounds the row ranges)
    if (t < n_tris) {
        double vs[3][3];
        bool finite = true;
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                vs[c][k] = tris[9 * t + 3 * c + k];
                finite &= isfinite(vs[c][k]);
            }
        if (!finite) atomicOr(err, (unsigned long long)kErrKeyRange);
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
        uint64_t nu = 0;
        if (finite && ana != 0.0) {  // else degenerate (or rejected): no rows
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
            const double lo_u = fmin(fmin(vs[0][u], vs[1][u]), vs[2][u]),
                         hi_u = fmax(fmax(vs[0][u], vs[1][u]), vs[2][u]);
            const double lo_w = fmin(fmin(vs[0][w], vs[1][w]), vs[2][w]),
                         hi_w = fmax(fmax(vs[0][w], vs[1][w]), vs[2][w]);
            const int64_t u0 = max(int64_t(0), int64_t(floor(lo_u)) - 1), u1 = min(side - 1, int64_t(floor(hi_u)) + 1);
            const int64_t w0 = max(int64_t(0), int64_t(floor(lo_w)) - 1), w1 = min(side - 1, int64_t(floor(hi_w)) + 1);
            S.nd = S.n[0] * vs[0][0] + S.n[1] * vs[0][1] + S.n[2] * vs[0][2];
            S.u0 = u0;
            S.w0 = w0;
            S.nw = w1 >= w0 ? uint32_t(w1 - w0 + 1) : 0u;
            nu = u1 >= u0 && S.nw ? uint64_t(u1 - u0 + 1) : 0u;
        }
        setup[t] = S;
        rows[t] = uint32_t(nu);  // at most the grid side, 2^21
        rect = nu * S.nw;
    }
    const uint64_t sum = block_reduce_sum<kVoxThreads>(rect, s_red);
    if (threadIdx.x == 0 && sum) atomicAdd(rect_columns, (unsigned long long)sum);
}

// Owner of global item g: the last i in [lo, hi) with off[i] <= g.
__device__ __forceinline__ uint64_t owner_of(const uint64_t* __restrict__ off, uint64_t lo, uint64_t hi, uint64_t g)
{
    while (hi - lo > 1) {
        const uint64_t mid = (lo + hi) / 2;
        if (off[mid] <= g) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Row r of the triangles' rectangles (row u = u0 + j of triangle t, j = r -
// row_off[t]): the columns [lo, hi] of the rectangle whose centres can pass
// the edge tests of the dominant projection (the plane of u and w).  Edge e
// passes where (nx * cu + ny * cw) + d >= 0; for ny != 0 that bounds cw on
// one side at c* = -(nx * cu + d) / ny.  The bound is widened by the rounding
// error of the test and of c* (far below 2^-46 of their magnitudes) plus one
// column, so it never drops a column the test keeps.
// Rows are handled in tiles of kVoxTile; each tile's column total goes to
// tile_cols for the scan that k_row_offsets finishes (a total of 2^32 or more
// sets kErrInternal in *err).
__global__ __launch_bounds__(kVoxThreads) void k_row_setup(const TriSetup* __restrict__ setup,
                                                           const uint64_t* __restrict__ row_off,
                                                           const uint64_t* __restrict__ n_tris_ptr,
                                                           const uint64_t* __restrict__ n_rows_ptr,
                                                           uint32_t* __restrict__ row_tri, uint32_t* __restrict__ row_u,
                                                           uint32_t* __restrict__ row_w0,
                                                           uint32_t* __restrict__ row_cols,
                                                           uint32_t* __restrict__ tile_cols,
                                                           unsigned long long* __restrict__ err)
{
    __shared__ uint64_t s_red[kVoxThreads / kWave];
    const uint64_t n_tris = *n_tris_ptr, n_rows = *n_rows_ptr;
    const uint64_t ntiles = (n_rows + kVoxTile - 1) / kVoxTile;
    for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        uint64_t tile_sum = 0;
        const uint64_t end = min(n_rows, (tile + 1) * kVoxTile);
        for (uint64_t r = tile * kVoxTile + threadIdx.x; r < end; r += kVoxThreads) {
            const uint64_t t = owner_of(row_off, 0, n_tris, r);
            const TriSetup& S = setup[t];
            const int64_t iu = S.u0 + int64_t(r - row_off[t]);
            const double cu = double(iu) + 0.5;
            int64_t lo = S.w0, hi = S.w0 + int64_t(S.nw) - 1;
            const double cw_far = double(hi) + 0.5;
            const int p = S.u;
            for (int e = 0; e < 3; ++e) {
                const double nx = S.ex[p][e], ny = S.ey[p][e], d = S.ed[p][e];
                const double g = nx * cu + d;
                if (ny == 0.0) {
                    // The test is g >= 0 for every column (adding ny * cw =
                    // +-0 changes no sign).
                    if (g < 0) hi = lo - 1;
                    continue;
                }
                const double cstar = -g / ny;
                const double mag = fabs(nx * cu) + fabs(ny) * cw_far + fabs(d);
                const double slack = mag * 0x1p-46 / fabs(ny) + fabs(cstar) * 0x1p-46 + 1.0;
                if (!(fabs(cstar) < 0x1p40) || !(slack < 0x1p30)) continue;  // no usable bound
                if (ny > 0) lo = max(lo, int64_t(floor(cstar - 0.5 - slack)));
                else hi = min(hi, int64_t(ceil(cstar - 0.5 + slack)));
            }
            const uint32_t cols = hi >= lo ? uint32_t(hi - lo + 1) : 0u;
            row_tri[r] = uint32_t(t);
            row_u[r] = uint32_t(iu);
            row_w0[r] = uint32_t(lo);
            row_cols[r] = cols;
            tile_sum += cols;
        }
        const uint64_t sum = block_reduce_sum<kVoxThreads>(tile_sum, s_red);
        if (threadIdx.x == 0) {
            tile_cols[tile] = uint32_t(sum);
            if (sum > 0xFFFFFFFFull) atomicOr(err, (unsigned long long)kErrInternal);
        }
    }
}

// col_off[r] = tile_off[tile] + the columns of the tile's rows before r.
__global__ __launch_bounds__(kVoxThreads) void k_row_offsets(const uint32_t* __restrict__ row_cols,
                                                             const uint64_t* __restrict__ tile_off,
                                                             const uint64_t* __restrict__ n_rows_ptr,
                                                             uint64_t* __restrict__ col_off)
{
    __shared__ uint32_t s_scan[kVoxThreads / kWave];
    const uint64_t n_rows = *n_rows_ptr;
    const uint64_t ntiles = (n_rows + kVoxTile - 1) / kVoxTile;
    for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const uint64_t r0 = tile * kVoxTile + uint64_t(threadIdx.x) * kVoxIpt;
        uint32_t c[kVoxIpt];
        uint32_t mine = 0;
#pragma unroll
        for (int i = 0; i < kVoxIpt; ++i) {
            c[i] = r0 + i < n_rows ? row_cols[r0 + i] : 0u;
            mine += c[i];
        }
        uint32_t unused;
        uint64_t o = tile_off[tile] + block_exclusive_scan<kVoxThreads>(mine, s_scan, unused);
#pragma unroll
        for (int i = 0; i < kVoxIpt; ++i) {
            if (r0 + i < n_rows) col_off[r0 + i] = o;
            o += c[i];
        }
    }
}

// Voxels of column (iu, iw) of triangle S (at most two), appended to k[] at
// `found`; returns the new count.
template <int N>
__device__ __forceinline__ int column_voxels(const TriSetup& S, int64_t iu, int64_t iw, int64_t side, uint64_t (&k)[N],
                                             int found)
{
    const double cu = double(iu) + 0.5, cw = double(iw) + 0.5;
    // The dominant projection (the plane of u and w) sees only the column:
    // its edge tests are the same for every voxel of it.
    for (int e = 0; e < 3; ++e)
        if (S.ex[S.u][e] * cu + S.ey[S.u][e] * cw + S.ed[S.u][e] < 0) return found;
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
            if (!((S.planes >> p) & 1) || p == S.u) continue;  // edge-on here, or tested above
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

// The rows of tile t's columns lie in [tile_row[t], tile_row[t + 1] + 1)
// (tile_row[t] = row of the tile's first column, one entry past the last
// tile), found by one thread per tile.
__global__ __launch_bounds__(kVoxThreads) void k_tile_rows(const uint64_t* __restrict__ col_off,
                                                           const uint64_t* __restrict__ n_rows_ptr,
                                                           const uint64_t* __restrict__ total_ptr,
                                                           uint64_t* __restrict__ tile_row)
{
    const uint64_t n_rows = *n_rows_ptr, total = *total_ptr;
    const uint64_t ntiles = (total + kVoxTile - 1) / kVoxTile;
    for (uint64_t t = uint64_t(blockIdx.x) * kVoxThreads + threadIdx.x; t <= ntiles;
         t += uint64_t(gridDim.x) * kVoxThreads)
        tile_row[t] = t < ntiles ? owner_of(col_off, 0, n_rows, t * kVoxTile) : n_rows - 1;
}

struct RowTable {
    const uint32_t* tri;
    const uint32_t* u;
    const uint32_t* w0;
    const uint64_t* col_off;  // exclusive scan of the rows' column counts
};

// Keys of the thread's kVoxIpt consecutive columns of tile `tile` (at most two
// per column) into k[]; returns how many.  One row search per thread; the
// columns then step along w, to the next non-empty row at a row's end.
constexpr int kVoxKeys = 2 * kVoxIpt;

__device__ __forceinline__ int thread_columns(const TriSetup* __restrict__ setup, const RowTable& R,
                                              const uint64_t* __restrict__ tile_row, uint64_t n_rows, uint64_t total,
                                              uint64_t tile, int64_t side, uint64_t (&k)[kVoxKeys])
{
    const uint64_t g0 = tile * kVoxTile + uint64_t(threadIdx.x) * kVoxIpt;
    if (g0 >= total) return 0;
    const uint64_t g1 = min(g0 + uint64_t(kVoxIpt), total);
    const uint64_t hi = tile_row[tile + 1] + 1;
    uint64_t r = owner_of(R.col_off, tile_row[tile], hi, g0);
    uint64_t r_end = r + 1 < n_rows ? R.col_off[r + 1] : total;
    uint32_t t = R.tri[r];
    int64_t iu = R.u[r], iw = int64_t(R.w0[r]) + int64_t(g0 - R.col_off[r]);
    int found = 0;
#pragma unroll 1
    for (uint64_t g = g0; g < g1; ++g) {
        if (g >= r_end) {  // the next row with columns starts at g
            r = owner_of(R.col_off, r + 1, hi, g);
            r_end = r + 1 < n_rows ? R.col_off[r + 1] : total;
            t = R.tri[r];
            iu = R.u[r];
            iw = R.w0[r];
        }
        found = column_voxels(setup[t], iu, iw, side, k, found);
        ++iw;
    }
    return found;
}

__global__ __launch_bounds__(kVoxThreads) void k_vox_count(const TriSetup* __restrict__ setup, RowTable R,
                                                           const uint64_t* __restrict__ tile_row,
                                                           const uint64_t* __restrict__ n_rows_ptr,
                                                           const uint64_t* __restrict__ total_ptr, int64_t side,
                                                           uint32_t* __restrict__ counts)
{
    __shared__ uint64_t s_red[kVoxThreads / kWave];
    const uint64_t n_rows = *n_rows_ptr, total = *total_ptr;
    const uint64_t ntiles = (total + kVoxTile - 1) / kVoxTile;
    for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        uint64_t k[kVoxKeys];
        const int mine = thread_columns(setup, R, tile_row, n_rows, total, tile, side, k);
        const uint64_t sum = block_reduce_sum<kVoxThreads>(uint64_t(mine), s_red);
        if (threadIdx.x == 0) counts[tile] = uint32_t(sum);
    }
}

__global__ __launch_bounds__(kVoxThreads) void k_vox_write(const TriSetup* __restrict__ setup, RowTable R,
                                                           const uint64_t* __restrict__ tile_row,
                                                           const uint64_t* __restrict__ n_rows_ptr,
                                                           const uint64_t* __restrict__ total_ptr, int64_t side,
                                                           const uint64_t* __restrict__ tile_offsets,
                                                           uint64_t* __restrict__ keys)
{
    __shared__ uint32_t s_scan[kVoxThreads / kWave];
    const uint64_t n_rows = *n_rows_ptr, total = *total_ptr;
    const uint64_t ntiles = (total + kVoxTile - 1) / kVoxTile;
    for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        uint64_t k[kVoxKeys];
        const int mine = thread_columns(setup, R, tile_row, n_rows, total, tile, side, k);
        uint32_t unused;
        const uint64_t o = tile_offsets[tile] + block_exclusive_scan<kVoxThreads>(uint32_t(mine), s_scan, unused);
#pragma unroll
        for (int j = 0; j < kVoxKeys; ++j)
            if (j < mine) keys[o + j] = k[j];
    }
}

// As k_vox_write, but the keys go to the tile's staging region
// stage[tile * 2 * kVoxTile ...] and the tile's count to counts[tile].
__global__ __launch_bounds__(kVoxThreads) void k_vox_emit(const TriSetup* __restrict__ setup, RowTable R,
                                                          const uint64_t* __restrict__ tile_row,
                                                          const uint64_t* __restrict__ n_rows_ptr,
                                                          const uint64_t* __restrict__ total_ptr, int64_t side,
                                                          uint32_t* __restrict__ counts, uint64_t* __restrict__ stage)
{
    __shared__ uint32_t s_scan[kVoxThreads / kWave];
    const uint64_t n_rows = *n_rows_ptr, total = *total_ptr;
    const uint64_t ntiles = (total + kVoxTile - 1) / kVoxTile;
    for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        uint64_t k[kVoxKeys];
        const int mine = thread_columns(setup, R, tile_row, n_rows, total, tile, side, k);
        uint32_t tile_keys;
        const uint64_t o = tile * 2 * kVoxTile + block_exclusive_scan<kVoxThreads>(uint32_t(mine), s_scan, tile_keys);
#pragma unroll
        for (int j = 0; j < kVoxKeys; ++j)
            if (j < mine) stage[o + j] = k[j];
        if (threadIdx.x == 0) counts[tile] = tile_keys;
    }
}

// keys[offsets[tile] + j] = stage[tile * 2 * kVoxTile + j] for j < counts[tile].
__global__ __launch_bounds__(kVoxThreads) void k_vox_compact(const uint64_t* __restrict__ stage,
                                                             const uint32_t* __restrict__ counts,
                                                             const uint64_t* __restrict__ offsets,
                                                             const uint64_t* __restrict__ total_ptr,
                                                             uint64_t* __restrict__ keys)
{
    const uint64_t ntiles = (*total_ptr + kVoxTile - 1) / kVoxTile;
    for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const uint32_t c = counts[tile];
        const uint64_t* src = stage + tile * 2 * kVoxTile;
        uint64_t* dst = keys + offsets[tile];
        for (uint32_t j = threadIdx.x; j < c; j += kVoxThreads) dst[j] = src[j];
    }
}

}  // namespace

static int vox_grid(uint64_t columns_bound)
{
    const uint64_t tiles = (columns_bound + kVoxTile - 1) / kVoxTile;
    return int(std::max<uint64_t>(1, std::min<uint64_t>(tiles, uint64_t(device_cu_count()) * 8)));
}

void launch_vox_emit(hipStream_t s, const TriSetup* setup, const VoxRows& rows, const uint64_t* n_rows_dev,
                     const uint64_t* total_cols_dev, uint64_t total_cols, int depth, uint64_t* tile_row,
                     uint32_t* tile_counts, uint64_t* stage)
{
    TimedLaunch t(kKVoxCount, s);
    const uint64_t tiles = (total_cols + kVoxTile - 1) / kVoxTile;
    hipLaunchKernelGGL(k_tile_rows, dim3(uint32_t((tiles + kVoxThreads) / kVoxThreads)), dim3(kVoxThreads), 0, s,
                       rows.col_off, n_rows_dev, total_cols_dev, tile_row);
    const RowTable R{rows.tri, rows.u, rows.w0, rows.col_off};
    hipLaunchKernelGGL(k_vox_emit, dim3(vox_grid(total_cols)), dim3(kVoxThreads), 0, s, setup, R, tile_row,
                       n_rows_dev, total_cols_dev, int64_t(1) << depth, tile_counts, stage);
}

void launch_vox_compact(hipStream_t s, const uint64_t* stage, const uint32_t* tile_counts, const uint64_t* tile_offsets,
                        const uint64_t* total_cols_dev, uint64_t total_cols, uint64_t* keys)
{
    TimedLaunch t(kKVoxWrite, s);
    hipLaunchKernelGGL(k_vox_compact, dim3(vox_grid(total_cols)), dim3(kVoxThreads), 0, s, stage, tile_counts,
                       tile_offsets, total_cols_dev, keys);
}

void launch_tri_setup(hipStream_t s, const double* tris, uint64_t n_tris, int depth, TriSetup* setup, uint32_t* rows,
                      uint64_t* rect_columns, unsigned long long* err)
{
    if (n_tris == 0) return;
    TimedLaunch t(kKVoxSetup, s);
    hipLaunchKernelGGL(k_tri_setup, dim3(uint32_t((n_tris + kVoxThreads - 1) / kVoxThreads)), dim3(kVoxThreads), 0, s,
                       tris, n_tris, int64_t(1) << depth, setup, rows,
                       reinterpret_cast<unsigned long long*>(rect_columns), err);
}

void launch_row_setup(hipStream_t s, const TriSetup* setup, const uint64_t* row_off, const uint64_t* n_tris_dev,
                      const uint64_t* n_rows_dev, uint64_t n_rows, const VoxRows& rows, uint32_t* tile_cols,
                      uint64_t* tile_off, uint64_t* total_cols_dev, unsigned long long* err)
{
    if (n_rows == 0) return;
    const uint32_t grid = uint32_t(std::min<uint64_t>((n_rows + kVoxTile - 1) / kVoxTile, device_cu_count() * 8ull));
    {
        TimedLaunch t(kKVoxSetup, s);
        hipLaunchKernelGGL(k_row_setup, dim3(grid), dim3(kVoxThreads), 0, s, setup, row_off, n_tris_dev, n_rows_dev,
                           rows.tri, rows.u, rows.w0, rows.cols, tile_cols, err);
    }
    launch_scan_counts(s, tile_cols, tile_off, n_rows_dev, kVoxTile, total_cols_dev);
    TimedLaunch t(kKVoxSetup, s);
    hipLaunchKernelGGL(k_row_offsets, dim3(grid), dim3(kVoxThreads), 0, s, rows.cols, tile_off, n_rows_dev,
                       rows.col_off);
}

void launch_vox_count(hipStream_t s, const TriSetup* setup, const VoxRows& rows, const uint64_t* n_rows_dev,
                      const uint64_t* total_cols_dev, uint64_t total_cols_bound, int depth, uint64_t* tile_row,
                      uint32_t* tile_counts)
{
    TimedLaunch t(kKVoxCount, s);
    const uint64_t tiles = (total_cols_bound + kVoxTile - 1) / kVoxTile;
    hipLaunchKernelGGL(k_tile_rows, dim3(uint32_t((tiles + kVoxThreads) / kVoxThreads)), dim3(kVoxThreads), 0, s,
                       rows.col_off, n_rows_dev, total_cols_dev, tile_row);
    const RowTable R{rows.tri, rows.u, rows.w0, rows.col_off};
    hipLaunchKernelGGL(k_vox_count, dim3(vox_grid(total_cols_bound)), dim3(kVoxThreads), 0, s, setup, R, tile_row,
                       n_rows_dev, total_cols_dev, int64_t(1) << depth, tile_counts);
}

void launch_vox_write(hipStream_t s, const TriSetup* setup, const VoxRows& rows, const uint64_t* tile_row,
                      const uint64_t* n_rows_dev, const uint64_t* total_cols_dev, uint64_t total_cols_bound, int depth,
                      const uint64_t* tile_offsets, uint64_t* keys)
{
    TimedLaunch t(kKVoxWrite, s);
    const RowTable R{rows.tri, rows.u, rows.w0, rows.col_off};
    hipLaunchKernelGGL(k_vox_write, dim3(vox_grid(total_cols_bound)), dim3(kVoxThreads), 0, s, setup, R, tile_row,
                       n_rows_dev, total_cols_dev, int64_t(1) << depth, tile_offsets, keys);
}

}  // namespace svdag
