// svdag_gen.cpp -- seeded synthetic Morton-key producers for the DAG build.
//
// The reference produces its input with an OpenGL conservative rasteriser
// (src/Voxelizer.cpp:254-310, Shaders/FragmentShader.glsl:19-32, 208-209):
// an unsorted uint64 list of Morton codes with duplicates.  These host
// generators produce lists with the same contract from seeded geometry:
//   * sphere shells          |‖v + ½ − c‖ − r| < ½, clipped to the grid;
//   * triangle surfaces      thin (26-separating style) surface voxelisation
//                            after Schwarz & Seidel 2010: one voxel per
//                            column along the dominant normal axis, kept when
//                            the voxel centre passes the three projected edge
//                            tests dilated by ½·max(|n_e.x|, |n_e.y|);
//   * value noise            trilinear lattice noise thresholded at ½;
//   * chunks                 a seeded mix of the above for one sub-volume.
// All floating-point arithmetic here is evaluated as written (the library builds
// this file with -ffp-contract=off): the device voxeliser repeats the triangle
// arithmetic operation for operation and must emit the same voxels.
// Duplicates are appended by resampling the unique set, then the list is
// Fisher-Yates shuffled.  Every random draw comes from xoshiro256** seeded by
// splitmix64 of (seed, item index), so output is independent of thread count.
//
// Morton layout (src/DAG_Compression_Fast.cpp:223-236, src/Core.h:252):
// bit 3k+2 = x bit k, bit 3k+1 = y bit k, bit 3k = z bit k.
#include "svdag.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <thread>
#include <vector>

namespace {

inline uint64_t splitmix64(uint64_t& x)
{
    uint64_t z = (x += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

struct Rng {
    uint64_t s[4];
    Rng(uint64_t seed, uint64_t stream)
    {
        uint64_t x = seed ^ (stream * 0xd1b54a32d192ed03ull);
        for (auto& v : s) v = splitmix64(x);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next()
    {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return double(next() >> 11) * 0x1.0p-53; }  // [0, 1)
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    uint64_t below(uint64_t n) { return uint64_t((unsigned __int128)next() * n >> 64); }
    double normal()
    {
        const double u1 = 1.0 - uniform(), u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

inline uint64_t split_by_3(uint32_t a)
{
    uint64_t x = a & 0x1fffff;
    x = (x | x << 32) & 0x1f00000000ffffull;
    x = (x | x << 16) & 0x1f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}

inline uint64_t morton(uint32_t x, uint32_t y, uint32_t z)
{
    return split_by_3(x) << 2 | split_by_3(y) << 1 | split_by_3(z);
}

void parallel_for(int threads, uint64_t n, const std::function<void(uint64_t, uint64_t, int)>& fn)
{
    threads = std::max(1, threads);
    if (threads == 1 || n < 2) {
        fn(0, n, 0);
        return;
    }
    std::vector<std::thread> pool;
    const uint64_t chunk = (n + threads - 1) / threads;
    for (int t = 0; t < threads; ++t) {
        const uint64_t b = std::min<uint64_t>(n, t * chunk), e = std::min<uint64_t>(n, b + chunk);
        if (b >= e) break;
        pool.emplace_back([&fn, b, e, t] { fn(b, e, t); });
    }
    for (auto& th : pool) th.join();
}

void sort_unique(std::vector<uint64_t>& v, int threads)
{
    const uint64_t n = v.size();
    int parts = 1;
    while (parts < std::max(1, threads)) parts <<= 1;
    if (n < (1u << 16)) parts = 1;
    std::vector<uint64_t> b(parts + 1);
    for (int i = 0; i <= parts; ++i) b[i] = n * i / parts;
    parallel_for(parts, parts, [&](uint64_t pb, uint64_t pe, int) {
        for (uint64_t p = pb; p < pe; ++p) std::sort(v.begin() + b[p], v.begin() + b[p + 1]);
    });
    std::vector<uint64_t> tmp(parts > 1 ? n : 0);
    std::vector<uint64_t>* src = &v;
    std::vector<uint64_t>* dst = &tmp;
    for (int w = 1; w < parts; w <<= 1) {
        std::vector<std::thread> pool;
        for (int i = 0; i < parts; i += 2 * w) {
            const uint64_t lo = b[i], mid = b[std::min(parts, i + w)], hi = b[std::min(parts, i + 2 * w)];
            pool.emplace_back([=] { std::merge(src->begin() + lo, src->begin() + mid, src->begin() + mid, src->begin() + hi, dst->begin() + lo); });
        }
        for (auto& th : pool) th.join();
        std::swap(src, dst);
    }
    if (src != &v) v.swap(tmp);
    v.erase(std::unique(v.begin(), v.end()), v.end());
}

using Emit = std::function<void(int64_t, int64_t, int64_t)>;

// Sphere shell: voxels whose centre is within ½ of the sphere surface.
void sphere_shell(double cx, double cy, double cz, double r, int64_t side, const Emit& emit)
{
    const int64_t x0 = std::max<int64_t>(0, int64_t(std::floor(cx - r - 1))), x1 = std::min<int64_t>(side - 1, int64_t(std::ceil(cx + r + 1)));
    const int64_t y0 = std::max<int64_t>(0, int64_t(std::floor(cy - r - 1))), y1 = std::min<int64_t>(side - 1, int64_t(std::ceil(cy + r + 1)));
    const int64_t z0 = std::max<int64_t>(0, int64_t(std::floor(cz - r - 1))), z1 = std::min<int64_t>(side - 1, int64_t(std::ceil(cz + r + 1)));
    for (int64_t x = x0; x <= x1; ++x)
        for (int64_t y = y0; y <= y1; ++y)
            for (int64_t z = z0; z <= z1; ++z) {
                const double dx = x + 0.5 - cx, dy = y + 0.5 - cy, dz = z + 0.5 - cz;
                const double d = std::sqrt(dx * dx + dy * dy + dz * dz);
                if (std::fabs(d - r) < 0.5) emit(x, y, z);
            }
}

struct V3 {
    double x, y, z;
    double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

// Thin surface voxelisation of one triangle (see file header).
void triangle_surface(V3 v0, V3 v1, V3 v2, int64_t side, const Emit& emit)
{
    const V3 n = cross(sub(v1, v0), sub(v2, v0));
    const double an[3] = {std::fabs(n.x), std::fabs(n.y), std::fabs(n.z)};
    const int a = an[0] >= an[1] && an[0] >= an[2] ? 0 : (an[1] >= an[2] ? 1 : 2);
    if (an[a] == 0.0) return;  // degenerate
    const int u = (a + 1) % 3, w = (a + 2) % 3;
    const V3 vs[3] = {v0, v1, v2};

    // Projected edge functions for the three axis planes (xy, yz, zx).
    struct Edge { double nx, ny, d; };
    Edge edges[3][3];
    for (int p = 0; p < 3; ++p) {
        const int i0 = p, i1 = (p + 1) % 3;  // plane (i0, i1); normal axis (p + 2) % 3
        const double sgn = n[(p + 2) % 3] >= 0 ? 1.0 : -1.0;
        for (int e = 0; e < 3; ++e) {
            const V3 a0 = vs[e], a1 = vs[(e + 1) % 3];
            const double ex = a1[i0] - a0[i0], ey = a1[i1] - a0[i1];
            Edge E;
            E.nx = -ey * sgn;
            E.ny = ex * sgn;
            E.d = -(E.nx * a0[i0] + E.ny * a0[i1]) + 0.5 * std::max(std::fabs(E.nx), std::fabs(E.ny));
            edges[p][e] = E;
        }
    }
    auto inside = [&](const double c[3]) {
        for (int p = 0; p < 3; ++p) {
            const int i0 = p, i1 = (p + 1) % 3;
            if (n[(p + 2) % 3] == 0.0) continue;  // triangle is edge-on in this projection
            for (int e = 0; e < 3; ++e) {
                const Edge& E = edges[p][e];
                if (E.nx * c[i0] + E.ny * c[i1] + E.d < 0) return false;
            }
        }
        return true;
    };

    const double lo_u = std::min({v0[u], v1[u], v2[u]}), hi_u = std::max({v0[u], v1[u], v2[u]});
    const double lo_w = std::min({v0[w], v1[w], v2[w]}), hi_w = std::max({v0[w], v1[w], v2[w]});
    const int64_t u0 = std::max<int64_t>(0, int64_t(std::floor(lo_u)) - 1), u1 = std::min<int64_t>(side - 1, int64_t(std::floor(hi_u)) + 1);
    const int64_t w0 = std::max<int64_t>(0, int64_t(std::floor(lo_w)) - 1), w1 = std::min<int64_t>(side - 1, int64_t(std::floor(hi_w)) + 1);
    const double nd = n.x * v0.x + n.y * v0.y + n.z * v0.z;
    for (int64_t iu = u0; iu <= u1; ++iu)
        for (int64_t iw = w0; iw <= w1; ++iw) {
            const double cu = iu + 0.5, cw = iw + 0.5;
            // Plane crossing along the dominant axis; the slab |n.(c - v0)| <= |n_a|/2
            // holds for the voxel k with k + ½ within ½ of it.
            const double ac = (nd - n[u] * cu - n[w] * cw) / n[a];
            const int64_t k0 = int64_t(std::floor(ac - 1.0)), k1 = int64_t(std::floor(ac));
            for (int64_t k = std::max<int64_t>(k0, 0); k <= std::min<int64_t>(k1, side - 1); ++k) {
                if (std::fabs(k + 0.5 - ac) > 0.5) continue;
                double c[3];
                c[a] = k + 0.5;
                c[u] = cu;
                c[w] = cw;
                if (!inside(c)) continue;
                int64_t xyz[3];
                xyz[a] = k;
                xyz[u] = iu;
                xyz[w] = iw;
                emit(xyz[0], xyz[1], xyz[2]);
            }
        }
}

inline double lognormal(Rng& r, double mu, double sigma) { return std::exp(mu + sigma * r.normal()); }

V3 random_direction(Rng& r)
{
    const double z = r.uniform(-1.0, 1.0), phi = r.uniform(0.0, 6.283185307179586);
    const double s = std::sqrt(std::max(0.0, 1.0 - z * z));
    return {s * std::cos(phi), s * std::sin(phi), z};
}

// Appends keys resampled from the unique set up to `n` in total, then shuffles.
void add_duplicates_to_and_shuffle(std::vector<uint64_t>& keys, uint64_t n, uint64_t seed)
{
    const uint64_t u = keys.size();
    if (u == 0) return;
    n = std::max(n, u);
    Rng r(seed, 0xD0B1E5ull);
    keys.reserve(n);
    for (uint64_t i = u; i < n; ++i) keys.push_back(keys[r.below(u)]);
    for (uint64_t i = n - 1; i > 0; --i) std::swap(keys[i], keys[r.below(i + 1)]);
}

// Duplicates make `dup_fraction` of the final list.
void add_duplicates_and_shuffle(std::vector<uint64_t>& keys, double dup_fraction, uint64_t seed)
{
    const double f = std::min(0.95, std::max(0.0, dup_fraction));
    add_duplicates_to_and_shuffle(keys, uint64_t(std::llround(double(keys.size()) / (1.0 - f))), seed);
}

int64_t side_of(int32_t depth) { return int64_t(1) << depth; }

std::vector<uint64_t> gen_spheres(uint64_t seed, int32_t depth, int32_t count, double rmin, double rmax, int threads)
{
    const int64_t side = side_of(depth);
    std::vector<std::vector<uint64_t>> local(std::max(1, threads));
    parallel_for(threads, uint64_t(count), [&](uint64_t b, uint64_t e, int t) {
        for (uint64_t i = b; i < e; ++i) {
            Rng r(seed, 1 + i);
            const double cx = r.uniform(0, double(side)), cy = r.uniform(0, double(side)), cz = r.uniform(0, double(side));
            const double rad = r.uniform(rmin, rmax);
            sphere_shell(cx, cy, cz, rad, side, [&](int64_t x, int64_t y, int64_t z) { local[t].push_back(morton(uint32_t(x), uint32_t(y), uint32_t(z))); });
        }
    });
    std::vector<uint64_t> all;
    for (auto& l : local) all.insert(all.end(), l.begin(), l.end());
    sort_unique(all, threads);
    return all;
}

// Triangle i of the triangle-surface workload: v0 uniform in the volume, the
// two other corners at lognormal distances in uniform directions.
void make_triangle(uint64_t seed, uint64_t i, int64_t side, double mu, double sigma, V3 (&v)[3])
{
    Rng r(seed, 1 + i);
    const V3 v0 = {r.uniform(0, double(side)), r.uniform(0, double(side)), r.uniform(0, double(side))};
    const double l1 = lognormal(r, mu, sigma), l2 = lognormal(r, mu, sigma);
    const V3 d1 = random_direction(r), d2 = random_direction(r);
    v[0] = v0;
    v[1] = {v0.x + l1 * d1.x, v0.y + l1 * d1.y, v0.z + l1 * d1.z};
    v[2] = {v0.x + l2 * d2.x, v0.y + l2 * d2.y, v0.z + l2 * d2.z};
}

std::vector<uint64_t> gen_triangles(uint64_t seed, int32_t depth, int32_t count, double mu, double sigma, int threads)
{
    const int64_t side = side_of(depth);
    std::vector<std::vector<uint64_t>> local(std::max(1, threads));
    // Interleave triangles over threads so heavy-tailed sizes balance.
    parallel_for(threads, uint64_t(std::max(1, threads)), [&](uint64_t tb, uint64_t te, int) {
        for (uint64_t t = tb; t < te; ++t)
            for (uint64_t i = t; i < uint64_t(count); i += uint64_t(std::max(1, threads))) {
                V3 v[3];
                make_triangle(seed, i, side, mu, sigma, v);
                triangle_surface(v[0], v[1], v[2], side, [&](int64_t x, int64_t y, int64_t z) { local[t].push_back(morton(uint32_t(x), uint32_t(y), uint32_t(z))); });
            }
    });
    std::vector<uint64_t> all;
    for (auto& l : local) all.insert(all.end(), l.begin(), l.end());
    sort_unique(all, threads);
    return all;
}

// Lattice value noise: value at lattice point (i,j,k) = top 24 bits of a hash, /2^24.
inline uint32_t lattice_value(uint64_t seed, int64_t i, int64_t j, int64_t k)
{
    uint64_t x = seed ^ (uint64_t(i) * 0x9E3779B97F4A7C15ull) ^ (uint64_t(j) * 0xC2B2AE3D27D4EB4Full) ^ (uint64_t(k) * 0x165667B19E3779F9ull);
    return uint32_t(splitmix64(x) >> 40);
}

// Occupancy of voxel (x,y,z): trilinear interpolation of lattice values at the
// voxel centre in 2^-24 fixed point, compared with 0.5 (= 2^23) exactly in integers.
inline bool noise_occupied(uint64_t seed, int64_t x, int64_t y, int64_t z, int64_t spacing)
{
    const int64_t i = x / spacing, j = y / spacing, k = z / spacing;
    // Centre offset within the cell in units of 1/(2*spacing): (2*(x mod s) + 1).
    const int64_t s2 = 2 * spacing;
    const int64_t fx = 2 * (x - i * spacing) + 1, fy = 2 * (y - j * spacing) + 1, fz = 2 * (z - k * spacing) + 1;
    __int128 acc = 0;
    for (int c = 0; c < 8; ++c) {
        const int64_t dx = c >> 2 & 1, dy = c >> 1 & 1, dz = c & 1;
        const int64_t wx = dx ? fx : s2 - fx, wy = dy ? fy : s2 - fy, wz = dz ? fz : s2 - fz;
        acc += (__int128)lattice_value(seed, i + dx, j + dy, k + dz) * (wx * wy * wz);
    }
    // value = acc / (s2^3 * 2^24) >= 1/2  <=>  acc * 2 >= s2^3 * 2^24
    return acc * 2 >= (__int128)(s2 * s2 * s2) << 24;
}

struct ChunkPlan {
    uint64_t n;          // total keys including duplicates
    uint64_t unique;     // distinct voxels
    double dup_fraction;
    int kind;            // 0 = sphere shells, 1 = triangle surface
};

ChunkPlan plan_chunk(uint64_t seed, uint64_t chunk, double median, double sigma, uint64_t nmin, uint64_t nmax)
{
    Rng r(seed, 0xC4A0C0000ull + chunk);
    ChunkPlan p;
    const double n = std::exp(std::log(median) + sigma * r.normal());
    p.n = uint64_t(std::llround(std::min(double(nmax), std::max(double(nmin), n))));
    p.dup_fraction = r.uniform(0.20, 0.30);
    p.unique = std::max<uint64_t>(1, uint64_t(std::llround(double(p.n) * (1.0 - p.dup_fraction))));
    p.kind = int(r.next() >> 63);
    return p;
}

// One chunk: whole shapes in generation order until `unique` distinct voxels,
// the last shape cut off in raster order; then duplicates up to n; shuffled.
std::vector<uint64_t> gen_chunk(uint64_t seed, uint64_t chunk, int32_t depth, const ChunkPlan& p)
{
    const int64_t side = side_of(depth);
    std::vector<uint64_t> bitmap((uint64_t(1) << (3 * depth)) / 64 + 1, 0);
    std::vector<uint64_t> keys;
    keys.reserve(p.n);
    const double scale = double(side) / 256.0;
    for (uint64_t shape = 0; keys.size() < p.unique; ++shape) {
        Rng r(seed ^ (chunk * 0x2545F4914F6CDD1Dull), 1 + shape);
        auto emit = [&](int64_t x, int64_t y, int64_t z) {
            if (keys.size() >= p.unique) return;
            const uint64_t c = morton(uint32_t(x), uint32_t(y), uint32_t(z));
            uint64_t& w = bitmap[c >> 6];
            const uint64_t bit = uint64_t(1) << (c & 63);
            if (w & bit) return;
            w |= bit;
            keys.push_back(c);
        };
        if (p.kind == 0) {
            const double cx = r.uniform(0, double(side)), cy = r.uniform(0, double(side)), cz = r.uniform(0, double(side));
            sphere_shell(cx, cy, cz, r.uniform(8.0 * scale, 40.0 * scale), side, emit);
        } else {
            const V3 v0 = {r.uniform(0, double(side)), r.uniform(0, double(side)), r.uniform(0, double(side))};
            const double l1 = lognormal(r, 4.0, 1.0), l2 = lognormal(r, 4.0, 1.0);
            const V3 d1 = random_direction(r), d2 = random_direction(r);
            triangle_surface(v0, {v0.x + l1 * d1.x, v0.y + l1 * d1.y, v0.z + l1 * d1.z},
                             {v0.x + l2 * d2.x, v0.y + l2 * d2.y, v0.z + l2 * d2.z}, side, emit);
        }
        if (shape > 100000000ull) break;  // cannot happen for the grid sizes used
    }
    add_duplicates_to_and_shuffle(keys, p.n, seed ^ (chunk * 0x9E3779B97F4A7C15ull));
    return keys;
}

template <class T>
int hand_out(const std::vector<T>& v, T** out, uint64_t* n)
{
    *n = v.size();
    *out = static_cast<T*>(std::malloc(std::max<size_t>(1, v.size() * sizeof(T))));
    if (!*out) return 1;
    if (!v.empty()) std::memcpy(*out, v.data(), v.size() * sizeof(T));
    return 0;
}

}  // namespace

extern "C" {

uint64_t svdag_morton_encode(uint32_t x, uint32_t y, uint32_t z) { return morton(x, y, z); }

int svdag_gen_sphere_shells(uint64_t seed, int32_t depth, int32_t count, double rmin, double rmax,
                            double dup_fraction, int32_t threads, uint64_t** keys, uint64_t* n)
{
    if (!keys || !n || depth < 1 || depth > 21 || count < 0) return 1;
    auto v = gen_spheres(seed, depth, count, rmin, rmax, threads);
    add_duplicates_and_shuffle(v, dup_fraction, seed);
    return hand_out(v, keys, n);
}

int svdag_gen_triangle_surfaces(uint64_t seed, int32_t depth, int32_t count, double mu, double sigma,
                                double dup_fraction, int32_t threads, uint64_t** keys, uint64_t* n)
{
    if (!keys || !n || depth < 1 || depth > 21 || count < 0) return 1;
    auto v = gen_triangles(seed, depth, count, mu, sigma, threads);
    add_duplicates_and_shuffle(v, dup_fraction, seed);
    return hand_out(v, keys, n);
}

int svdag_gen_triangles(uint64_t seed, int32_t depth, int32_t count, double mu, double sigma, double* out)
{
    if (depth < 1 || depth > 21 || count < 0 || (count && !out)) return SVDAG_ERR_ARGUMENT;
    const int64_t side = side_of(depth);
    for (int32_t i = 0; i < count; ++i) {
        V3 v[3];
        make_triangle(seed, uint64_t(i), side, mu, sigma, v);
        for (int c = 0; c < 3; ++c) {
            out[9 * i + 3 * c + 0] = v[c].x;
            out[9 * i + 3 * c + 1] = v[c].y;
            out[9 * i + 3 * c + 2] = v[c].z;
        }
    }
    return 0;
}

int svdag_gen_voxelize(const double* tris, uint64_t n_tris, int32_t depth, int32_t threads, uint64_t** keys,
                       uint64_t* n)
{
    if (depth < 1 || depth > 21 || !keys || !n || (n_tris && !tris)) return SVDAG_ERR_ARGUMENT;
    const int64_t side = side_of(depth);
    std::vector<std::vector<uint64_t>> local(std::max(1, threads));
    parallel_for(threads, n_tris, [&](uint64_t b, uint64_t e, int t) {
        for (uint64_t i = b; i < e; ++i) {
            const double* p = tris + 9 * i;
            triangle_surface({p[0], p[1], p[2]}, {p[3], p[4], p[5]}, {p[6], p[7], p[8]}, side,
                             [&](int64_t x, int64_t y, int64_t z) { local[t].push_back(morton(uint32_t(x), uint32_t(y), uint32_t(z))); });
        }
    });
    std::vector<uint64_t> all;
    for (auto& l : local) all.insert(all.end(), l.begin(), l.end());
    sort_unique(all, threads);
    return hand_out(all, keys, n);
}

int svdag_gen_value_noise(uint64_t seed, int32_t depth, int32_t spacing, uint32_t sub_depth, uint64_t sub_index,
                          int32_t threads, uint64_t** keys, uint64_t* n)
{
    // Voxels of the Morton-aligned sub-cube `sub_index` (side 2^sub_depth) of a
    // 2^depth grid, emitted in shuffled order with codes in the full grid.
    if (!keys || !n || depth < 1 || depth > 21 || int32_t(sub_depth) > depth || spacing < 1) return 1;
    const uint64_t sub_voxels = uint64_t(1) << (3 * sub_depth);
    const uint64_t base = sub_index << (3 * sub_depth);
    if ((sub_index >> (3 * (depth - int32_t(sub_depth)))) != 0) return 1;
    std::vector<std::vector<uint64_t>> local(std::max(1, threads));
    parallel_for(threads, sub_voxels / 64, [&](uint64_t b, uint64_t e, int t) {
        for (uint64_t blk = b; blk < e; ++blk)
            for (uint64_t j = 0; j < 64; ++j) {
                const uint64_t code = base + blk * 64 + j;
                uint32_t xyz[3] = {0, 0, 0};
                for (int bit = 0; bit < depth; ++bit)
                    for (int a = 0; a < 3; ++a) xyz[a] |= uint32_t((code >> (3 * bit + (2 - a))) & 1) << bit;
                if (noise_occupied(seed, xyz[0], xyz[1], xyz[2], spacing)) local[t].push_back(code);
            }
    });
    std::vector<uint64_t> all;
    for (auto& l : local) all.insert(all.end(), l.begin(), l.end());
    Rng r(seed ^ sub_index, 0x5A0FF1Eull);
    for (uint64_t i = all.size(); i > 1; --i) std::swap(all[i - 1], all[r.below(i)]);
    return hand_out(all, keys, n);
}

int svdag_gen_chunk_plan(uint64_t seed, uint64_t num_chunks, double median, double sigma, uint64_t nmin,
                         uint64_t nmax, uint64_t* n_out, uint64_t* unique_out, int32_t* kind_out)
{
    for (uint64_t c = 0; c < num_chunks; ++c) {
        const ChunkPlan p = plan_chunk(seed, c, median, sigma, nmin, nmax);
        if (n_out) n_out[c] = p.n;
        if (unique_out) unique_out[c] = p.unique;
        if (kind_out) kind_out[c] = p.kind;
    }
    return 0;
}

int svdag_gen_chunk(uint64_t seed, uint64_t chunk, int32_t depth, double median, double sigma, uint64_t nmin,
                    uint64_t nmax, uint64_t* keys, uint64_t capacity, uint64_t* n)
{
    // Writes the chunk's chunk-local keys (< 2^(3*depth)) into `keys`.
    if (!keys || !n || depth < 1 || depth > 12) return 1;
    const ChunkPlan p = plan_chunk(seed, chunk, median, sigma, nmin, nmax);
    if (p.n > capacity) return 2;
    const auto v = gen_chunk(seed, chunk, depth, p);
    std::memcpy(keys, v.data(), v.size() * 8);
    *n = v.size();
    return 0;
}

void svdag_gen_free(void* p) { std::free(p); }

}  // extern "C"
