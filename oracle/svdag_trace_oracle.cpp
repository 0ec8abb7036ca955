// svdag_trace_oracle.cpp -- CPU restatement of the reference's tracer walk.
//
// TEST INFRASTRUCTURE ONLY (see svdag_oracle.h).  Restates trace_paths
// (src/DAGTracer.cu:177-428) and the camera set-up it is driven with
// (FDAGTracer::TracePathsParams, src/DAGTracer.h:17-29) as a recursive
// depth-first descent: a node's crossed children are visited in the ray's
// front-to-back order and the first voxel reached ends the walk, which is the
// order the reference's explicit stack produces.  Built with
// -ffp-contract=off: every float32 / float64 operation below is one IEEE
// operation, in the reference's order.
#include "svdag_oracle.h"

#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>

namespace {

struct Walk {
    const uint32_t* words;
    uint64_t n_words;
    int depth;
    uint32_t debug_level;
    // ray
    float o[3], d[3], inv[3];
    unsigned order;
    // result
    uint32_t debug_index = 0;
    bool bad = false;
    uint32_t hit[3] = {0, 0, 0};
    // leaf occupancy of the current 4x4x4 block (src/DAGTracer.cu:95-138)
    uint64_t leaf = 0;
};

// (a > b) ? a : b and its mirror, as src/CudaMath.h:12-15.
float rmax(float a, float b) { return a > b ? a : b; }
float rmin(float a, float b) { return a < b ? a : b; }

// compute_intersection_mask (src/DAGTracer.cu:177-290).
unsigned intersection_mask(const Walk& w, bool is_root, int level, const uint32_t pos[3])
{
    const unsigned shift = unsigned(w.depth - level);
    const float radius = float(1u << (shift - 1));
    const float cx = radius + float(pos[0] << shift), cy = radius + float(pos[1] << shift),
                cz = radius + float(pos[2] << shift);
    const float rx = cx - w.o[0], ry = cy - w.o[1], rz = cz - w.o[2];
    const float tx = rx * w.inv[0], ty = ry * w.inv[1], tz = rz * w.inv[2];
    const float sx = radius * std::fabs(w.inv[0]), sy = radius * std::fabs(w.inv[1]),
                sz = radius * std::fabs(w.inv[2]);
    const float tmin = rmax(rmax(tx - sx, rmax(ty - sy, tz - sz)), 0.0f);
    const float tmax = rmin(tx + sx, rmin(ty + sy, tz + sz));
    if (is_root && tmin >= tmax) return 0;
    const float h = 0.5f * (tmin + tmax);
    const unsigned first = (h * w.d[0] >= rx ? 4u : 0u) + (h * w.d[1] >= ry ? 2u : 0u) + (h * w.d[2] >= rz ? 1u : 0u);
    unsigned mask = 1u << first;
    const float e = 1e-4f;
    if (tmin <= tx && tx <= tmax) {
        const float py = tx * w.d[1], pz = tx * w.d[2];
        const unsigned A = (py >= ry - e ? 0xCCu : 0u) | (py <= ry + e ? 0x33u : 0u);
        const unsigned B = (pz >= rz - e ? 0xAAu : 0u) | (pz <= rz + e ? 0x55u : 0u);
        mask |= A & B;
    }
    if (tmin <= ty && ty <= tmax) {
        const float px = ty * w.d[0], pz = ty * w.d[2];
        const unsigned C = (px >= rx - e ? 0xF0u : 0u) | (px <= rx + e ? 0x0Fu : 0u);
        const unsigned D = (pz >= rz - e ? 0xAAu : 0u) | (pz <= rz + e ? 0x55u : 0u);
        mask |= C & D;
    }
    if (tmin <= tz && tz <= tmax) {
        const float px = tz * w.d[0], py = tz * w.d[1];
        const unsigned E = (px >= rx - e ? 0xF0u : 0u) | (px <= rx + e ? 0x0Fu : 0u);
        const unsigned F = (py >= ry - e ? 0xCCu : 0u) | (py <= ry + e ? 0x33u : 0u);
        mask |= E & F;
    }
    return mask;
}

// Visits the node at `level` (grid position pos, word index `index`, child
// mask `child_mask`); returns true when the walk is over (a voxel was reached
// or the layout was left).  `index` is the last interior node's word index,
// kept through the two leaf levels as the reference keeps it.
bool visit(Walk& w, int level, const uint32_t pos[3], uint32_t index, unsigned child_mask, bool is_root)
{
    const int leaf_level = w.depth - 2;
    unsigned todo = child_mask & intersection_mask(w, is_root, level, pos);
    while (todo) {
        // next_child (src/DAGTracer.cu:31-41)
        unsigned child = 0;
        for (unsigned k = 0; k < 8; ++k)
            if (todo & (1u << (k ^ w.order))) {
                child = k ^ w.order;
                break;
            }
        todo &= ~(1u << child);
        const uint32_t cpos[3] = {pos[0] << 1 | (child >> 2 & 1u), pos[1] << 1 | (child >> 1 & 1u),
                                  pos[2] << 1 | (child & 1u)};
        const int clevel = level + 1;
        if (uint32_t(w.depth - 2 - clevel) == w.debug_level) w.debug_index = index;
        if (clevel == w.depth) {
            w.hit[0] = cpos[0];
            w.hit[1] = cpos[1];
            w.hit[2] = cpos[2];
            return true;
        }
        uint32_t cindex = index;
        unsigned cmask;
        if (clevel <= leaf_level) {
            const uint64_t at = uint64_t(index) + 1 + unsigned(__builtin_popcount(child_mask & ((1u << child) - 1u)));
            if (at >= w.n_words) {
                w.bad = true;
                return true;
            }
            const uint32_t target = w.words[at];
            if (clevel < leaf_level) {
                if (target >= w.n_words) {
                    w.bad = true;
                    return true;
                }
                cindex = target;
                cmask = w.words[target] & 0xFFu;
            } else {
                if (uint64_t(target) + 1 >= w.n_words) {
                    w.bad = true;
                    return true;
                }
                w.leaf = uint64_t(w.words[target + 1]) << 32 | w.words[target];
                cmask = 0;
                for (unsigned b = 0; b < 8; ++b)
                    if ((w.leaf >> (8 * b)) & 0xFFu) cmask |= 1u << b;
            }
        } else {
            cmask = unsigned(w.leaf >> (8 * child)) & 0xFFu;
        }
        if (visit(w, clevel, cpos, cindex, cmask, false)) return true;
    }
    return false;
}

uint32_t murmur32(uint32_t h)
{
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}

void trace_pixel(const uint32_t* words, uint64_t n_words, int depth, const ora_trace_params& p, uint32_t x,
                 uint32_t y, uint32_t* path_out, uint32_t* color_out)
{
    Walk w;
    w.words = words;
    w.n_words = n_words;
    w.depth = depth;
    w.debug_level = p.debug_level;
    // src/DAGTracer.cu:308-319
    double v[3];
    for (int a = 0; a < 3; ++a)
        v[a] = p.ray_min[a] + double(x) * p.ray_ddx[a] + double(y) * p.ray_ddy[a] - p.camera_position[a];
    const double len = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    const double s = 1 / len;
    for (int a = 0; a < 3; ++a) {
        w.o[a] = float(p.camera_position[a]);
        w.d[a] = float(s * v[a]);
        w.inv[a] = float(1. / double(w.d[a]));
    }
    w.order = (w.d[0] < 0.f ? 4u : 0u) + (w.d[1] < 0.f ? 2u : 0u) + (w.d[2] < 0.f ? 1u : 0u);
    const uint32_t origin[3] = {0, 0, 0};
    bool hit = false;
    if (n_words > 0) hit = visit(w, 0, origin, 0, words[0] & 0xFFu, true);
    uint32_t path[3] = {0, 0, 0};
    if (w.bad) path[0] = path[1] = path[2] = 0xFFFFFFFFu;
    else if (hit) std::copy(w.hit, w.hit + 3, path);
    if (path_out) std::copy(path, path + 3, path_out);
    if (color_out) {
        uint32_t c;
        if (w.bad) {
            c = 0;
        } else if (p.debug_level == 0xFFFFFFFFu) {
            // float3_to_rgb888 (src/DAGTracer.cu:20-28)
            uint32_t rgb = 0;
            for (int a = 0; a < 3; ++a) {
                const float f = float(path[a] & 0xFFu) / float(0xFF);
                const float q = std::fmin(1.0f, std::fmax(0.0f, f));
                rgb |= uint32_t(q * 255.0f) << (8 * a);
            }
            c = rgb;
        } else {
            c = murmur32(w.debug_index);
        }
        *color_out = c;
    }
}

}  // namespace

extern "C" int ora_trace_paths(const uint32_t* words, uint64_t n_words, int32_t depth, const ora_trace_params* p,
                               int32_t threads, uint32_t* paths, uint32_t* colors)
{
    if (!p || depth < 3 || depth > 21 || (n_words && !words)) return 1;
    const uint32_t W = p->width, H = p->height;
    const int T = std::max(1, std::min<int>(threads, int(H ? H : 1)));
    std::vector<std::thread> pool;
    for (int t = 0; t < T; ++t)
        pool.emplace_back([=]() {
            for (uint32_t y = uint32_t(t); y < H; y += uint32_t(T))
                for (uint32_t x = 0; x < W; ++x) {
                    const uint64_t i = uint64_t(y) * W + x;
                    trace_pixel(words, n_words, depth, *p, x, y, paths ? paths + 3 * i : nullptr,
                                colors ? colors + i : nullptr);
                }
        });
    for (auto& th : pool) th.join();
    return 0;
}
