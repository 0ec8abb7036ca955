/*
 * svdag_oracle.cpp -- CPU oracle for the sparse-voxel-DAG build.
 *
 * TEST INFRASTRUCTURE ONLY: the expected results the HIP library's tests
 * compare against, and the timed host-core baseline.  Restates, step for step, the reference algorithm of
 * Phyronnaz/DAG_Compression_Fast:
 *   CreateDAG        src/DAGCompression.cu:1003-1107
 *   SortLevel        src/DAGCompression.cu:864-914
 *   ExtractLevel     src/DAGCompression.cu:916-974
 *   ComputeHashes    src/DAGCompression.cu:827-833, 758-774
 *   MurmurHash128xN  src/Utils.h:119-154 (MurmurHash64 src/Utils.h:86-94)
 *   Merge/MergeDAGs  src/DAGCompression.cu:1109-1320, 1447-1548
 *   CreateFinalDAG   src/DAGCompression.cu:834-861, 1550-1610
 *
 * The GPU primitives the reference calls (CUB radix sort, select-unique,
 * reduce-by-key, scans; thrust scatter/transform; moderngpu merge) are
 * replaced by their textbook definitions (std::sort, run detection, prefix
 * sums, std::merge), parallelised over host threads with std::thread.
 *
 * Deviations from the reference, all outside its contract:
 *  - keys >= 2^(3*depth) are rejected with an error (the reference assumes
 *    them away and sorts only the low 3*depth bits);
 *  - no Fragments.Num()/2 under-allocation (DAGCompression.cu:1050-1051);
 *  - the input is never clobbered (DAGCompression.cu:1028-1033);
 *  - when two different nodes share a 64-bit hash, the reference scatter
 *    keeps an unspecified one; here the first in stable sorted order is kept
 *    and the event is counted in hash_collisions.
 */
#include "svdag_oracle.h"

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>
#include <thread>
#include <utility>
#include <vector>

namespace {

thread_local std::string g_error;

int fail(const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_error = buf;
    return 1;
}

constexpr uint32_t kEmpty = 0xFFFFFFFFu;

// ---------------------------------------------------------------------------
// Hash (src/Utils.h:41-44, 86-94, 119-154)
// ---------------------------------------------------------------------------
inline uint64_t rotl64(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }

inline uint64_t fmix64(uint64_t h)  // Utils::MurmurHash64, Utils.h:86-94
{
    h ^= h >> 33;
    h *= 0xff51afd7ed558ccdull;
    h ^= h >> 33;
    h *= 0xc4ceb9fe1a85ec53ull;
    h ^= h >> 33;
    return h;
}

void murmur128(const uint64_t* data, uint32_t size, uint32_t seed, uint64_t len_xor, uint64_t out[2])
{
    uint64_t h1 = seed, h2 = seed;
    const uint64_t c1 = 0x87c37b91114253d5ull;
    const uint64_t c2 = 0x4cf5ad432745937full;
    for (uint32_t i = 0; i < size; ++i) {
        uint64_t k1 = data[2 * i + 0];
        uint64_t k2 = data[2 * i + 1];
        k1 *= c1; k1 = rotl64(k1, 31); k1 *= c2; h1 ^= k1;
        h1 = rotl64(h1, 27); h1 += h2; h1 = h1 * 5 + 0x52dce729;
        k2 *= c2; k2 = rotl64(k2, 33); k2 *= c1; h2 ^= k2;
        h2 = rotl64(h2, 31); h2 += h1; h2 = h2 * 5 + 0x38495ab5;
    }
    h1 ^= len_xor;
    h2 ^= len_xor;
    h1 += h2;
    h2 += h1;
    h1 = fmix64(h1);
    h2 = fmix64(h2);
    h1 += h2;
    h2 += h1;
    out[0] = h1;
    out[1] = h2;
}

// HashChildrenHashes, DAGCompression.cu:758-763 (Size = 4 blocks, Seed = 0).
inline uint64_t hash_children(const uint64_t ch[8])
{
    uint64_t out[2];
    murmur128(ch, 4, 0, 4, out);
    return out[0] == 0 ? 1 : out[0];
}

// ---------------------------------------------------------------------------
// Host parallelism helpers
// ---------------------------------------------------------------------------
void parallel_for(int threads, uint64_t n, const std::function<void(uint64_t, uint64_t)>& fn)
{
    if (threads <= 1 || n < 4096) {
        fn(0, n);
        return;
    }
    std::vector<std::thread> pool;
    const uint64_t chunk = (n + threads - 1) / threads;
    for (int t = 0; t < threads; ++t) {
        const uint64_t b = std::min<uint64_t>(n, t * chunk), e = std::min<uint64_t>(n, b + chunk);
        if (b >= e) break;
        pool.emplace_back([&fn, b, e] { fn(b, e); });
    }
    for (auto& th : pool) th.join();
}

// Sort T chunks in parallel, then merge pairwise (parallel within a round).
template <class T>
void parallel_sort(std::vector<T>& v, int threads)
{
    const uint64_t n = v.size();
    if (threads <= 1 || n < (1u << 16)) {
        std::sort(v.begin(), v.end());
        return;
    }
    int parts = 1;
    while (parts < threads) parts <<= 1;
    std::vector<uint64_t> bounds(parts + 1);
    for (int i = 0; i <= parts; ++i) bounds[i] = n * i / parts;
    {
        std::vector<std::thread> pool;
        for (int i = 0; i < parts; ++i)
            pool.emplace_back([&, i] { std::sort(v.begin() + bounds[i], v.begin() + bounds[i + 1]); });
        for (auto& th : pool) th.join();
    }
    std::vector<T> tmp(n);
    std::vector<T>* src = &v;
    std::vector<T>* dst = &tmp;
    for (int width = 1; width < parts; width <<= 1) {
        std::vector<std::thread> pool;
        for (int i = 0; i < parts; i += 2 * width) {
            const uint64_t b = bounds[i], m = bounds[std::min(parts, i + width)], e = bounds[std::min(parts, i + 2 * width)];
            pool.emplace_back([=] {
                std::merge(src->begin() + b, src->begin() + m, src->begin() + m, src->begin() + e, dst->begin() + b);
            });
        }
        for (auto& th : pool) th.join();
        std::swap(src, dst);
    }
    if (src != &v) v.swap(tmp);
}

// Positions i where key(i) != key(i-1) (i == 0 always starts a run), in order.
std::vector<uint64_t> run_starts(uint64_t n, int threads, const std::function<uint64_t(uint64_t)>& key)
{
    const int parts = std::max(1, threads);
    std::vector<std::vector<uint64_t>> local(parts);
    const uint64_t chunk = (n + parts - 1) / parts;
    parallel_for(parts, parts, [&](uint64_t pb, uint64_t pe) {
        for (uint64_t p = pb; p < pe; ++p) {
            const uint64_t b = std::min(n, p * chunk), e = std::min(n, b + chunk);
            for (uint64_t i = b; i < e; ++i)
                if (i == 0 || key(i) != key(i - 1)) local[p].push_back(i);
        }
    });
    std::vector<uint64_t> out;
    for (auto& l : local) out.insert(out.end(), l.begin(), l.end());
    return out;
}

struct Level {
    std::vector<uint8_t> masks;
    std::vector<uint32_t> children;  // 8 per node
    std::vector<uint64_t> hashes;
};

struct Dag {
    std::vector<Level> levels;
    std::vector<uint64_t> leaves;
    uint64_t collisions = 0;
    uint64_t unique_keys = 0, leaf_positions = 0;
    uint64_t pre_dedup[32] = {};
};

// SortLevel, DAGCompression.cu:864-914.  Stable sort of (hash, index), unique
// rank by adjacent difference + inclusive sum, scatter to ascending-hash order.
// Returns the index -> unique-rank map; fills uniq (sorted unique hashes) and
// rep (the index whose content represents each unique hash).
void sort_level(const std::vector<uint64_t>& hashes, int threads, std::vector<uint32_t>& map,
                std::vector<uint64_t>& uniq, std::vector<uint32_t>& rep)
{
    const uint64_t m = hashes.size();
    std::vector<std::pair<uint64_t, uint32_t>> p(m);
    parallel_for(threads, m, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; ++i) p[i] = {hashes[i], uint32_t(i)};
    });
    parallel_sort(p, threads);
    map.assign(m, 0);
    uniq.clear();
    rep.clear();
    uint32_t rank = 0;
    for (uint64_t j = 0; j < m; ++j) {
        if (j > 0 && p[j].first != p[j - 1].first) ++rank;
        if (j == 0 || p[j].first != p[j - 1].first) {
            uniq.push_back(p[j].first);
            rep.push_back(p[j].second);
        }
        map[p[j].second] = rank;
    }
}

bool check_keys(const uint64_t* keys, uint64_t n, int32_t depth, int threads)
{
    if (depth >= 22) return true;
    const uint64_t limit = uint64_t(1) << (3 * depth);
    std::vector<char> bad(std::max(1, threads), 0);
    const uint64_t chunk = (n + bad.size() - 1) / bad.size();
    parallel_for(int(bad.size()), bad.size(), [&](uint64_t pb, uint64_t pe) {
        for (uint64_t t = pb; t < pe; ++t)
            for (uint64_t i = t * chunk; i < std::min(n, (t + 1) * chunk); ++i)
                if (keys[i] >= limit) { bad[t] = 1; break; }
    });
    for (char c : bad)
        if (c) return false;
    return true;
}

int create_dag(const uint64_t* in_keys, uint64_t n, int32_t depth, int threads, Dag& dag)
{
    if (depth < 3 || depth > 21) return fail("depth %d outside [3, 21]", depth);
    if (n >= (uint64_t(1) << 32)) return fail("key count %llu exceeds 2^32 - 1", (unsigned long long)n);
    if (!check_keys(in_keys, n, depth, threads))
        return fail("a key is >= 2^(3*depth) for depth %d", depth);
    threads = std::max(1, threads);

    dag.levels.assign(depth - 2, Level{});  // DAGCompression.cu:1008
    if (n == 0) return 0;                   // DAGCompression.cu:1010-1013

    // SortKeys + SelectUnique (DAGCompression.cu:1024-1046).
    std::vector<uint64_t> frags(in_keys, in_keys + n);
    parallel_sort(frags, threads);
    frags.erase(std::unique(frags.begin(), frags.end()), frags.end());
    dag.unique_keys = frags.size();

    // Leaf ReduceByKey: key code>>6, value 1<<(code&63), OR (DAGCompression.cu:1044-1068).
    std::vector<uint64_t> local_leaves, parents;
    {
        const auto starts = run_starts(frags.size(), threads, [&](uint64_t i) { return frags[i] >> 6; });
        local_leaves.assign(starts.size(), 0);
        parents.assign(starts.size(), 0);
        parallel_for(threads, starts.size(), [&](uint64_t b, uint64_t e) {
            for (uint64_t r = b; r < e; ++r) {
                const uint64_t s = starts[r], t = r + 1 < starts.size() ? starts[r + 1] : frags.size();
                uint64_t mask = 0;
                for (uint64_t i = s; i < t; ++i) mask |= uint64_t(1) << (frags[i] & 63);
                local_leaves[r] = mask;
                parents[r] = frags[s] >> 6;
            }
        });
    }
    dag.leaf_positions = parents.size();
    frags.swap(parents);

    // SortLevel(Leaves) (DAGCompression.cu:1070-1071): the hash of a leaf is its mask.
    std::vector<uint32_t> frag_to_child;
    std::vector<uint32_t> rep;
    sort_level(local_leaves, threads, frag_to_child, dag.leaves, rep);

    std::vector<uint64_t> prev_hashes = dag.leaves;
    for (int32_t level_depth = depth - 3; level_depth >= 0; --level_depth) {
        // ExtractLevel (DAGCompression.cu:916-974).
        const auto starts = run_starts(frags.size(), threads, [&](uint64_t i) { return frags[i] >> 3; });
        const uint64_t m = starts.size();
        dag.pre_dedup[level_depth] = m;
        std::vector<uint8_t> masks(m, 0);
        std::vector<uint32_t> children(8 * m, kEmpty);
        std::vector<uint64_t> parent_codes(m), hashes(m);
        parallel_for(threads, m, [&](uint64_t b, uint64_t e) {
            for (uint64_t p = b; p < e; ++p) {
                const uint64_t s = starts[p], t = p + 1 < m ? starts[p + 1] : frags.size();
                for (uint64_t i = s; i < t; ++i) {
                    const uint32_t slot = uint32_t(frags[i] & 7);
                    masks[p] |= uint8_t(1u << slot);
                    children[8 * p + slot] = frag_to_child[i];
                }
                parent_codes[p] = frags[s] >> 3;
                // ComputeHashes / HashChildren (DAGCompression.cu:765-774, 827-833).
                uint64_t ch[8];
                for (int k = 0; k < 8; ++k) {
                    const uint32_t c = children[8 * p + k];
                    ch[k] = c == kEmpty ? 0 : prev_hashes[c];
                }
                hashes[p] = hash_children(ch);
            }
        });
        frags.swap(parent_codes);

        // SortLevel(Level.Hashes, ..., sort ChildrenIndices and ChildMasks) (DAGCompression.cu:1087).
        Level& out = dag.levels[level_depth];
        sort_level(hashes, threads, frag_to_child, out.hashes, rep);
        const uint64_t q = out.hashes.size();
        out.masks.resize(q);
        out.children.resize(8 * q);
        for (uint64_t r = 0; r < q; ++r) {
            out.masks[r] = masks[rep[r]];
            std::memcpy(&out.children[8 * r], &children[8 * rep[r]], 32);
        }
        for (uint64_t p = 0; p < m; ++p) {
            const uint32_t r = frag_to_child[p];
            if (masks[p] != out.masks[r] || std::memcmp(&children[8 * p], &out.children[8 * r], 32) != 0)
                ++dag.collisions;
        }
        prev_hashes = out.hashes;
    }
    if (prev_hashes.size() != 1) return fail("root level holds %zu nodes", prev_hashes.size());  // :1093
    return 0;
}

// Merge (DAGCompression.cu:1109-1163): stable merge of two strictly ascending
// arrays (A before B on ties), adjacent difference + inclusive sum, scatter of
// the A / B indices to their merged unique rank.
uint64_t merge_sorted(const std::vector<uint64_t>& a, const std::vector<uint64_t>& b,
                      std::vector<uint32_t>& map_a, std::vector<uint32_t>& map_b, std::vector<uint64_t>& merged)
{
    map_a.assign(a.size(), 0);
    map_b.assign(b.size(), 0);
    merged.clear();
    uint64_t i = 0, j = 0;
    while (i < a.size() || j < b.size()) {
        const bool take_a = j >= b.size() || (i < a.size() && !(b[j] < a[i]));
        const uint64_t key = take_a ? a[i] : b[j];
        if (merged.empty() || merged.back() != key) merged.push_back(key);
        const uint32_t r = uint32_t(merged.size() - 1);
        if (take_a) map_a[i++] = r;
        else map_b[j++] = r;
    }
    return merged.size();
}

// MergeDAGs(Pool, A, B) (DAGCompression.cu:1165-1320).
Dag merge_pair(const Dag& a, const Dag& b)
{
    if (a.leaves.empty()) return b;
    if (b.leaves.empty()) return a;
    const int nl = int(a.levels.size());
    Dag out;
    out.levels.resize(nl);
    out.collisions = a.collisions + b.collisions;

    std::vector<uint32_t> map_a, map_b, prev_a, prev_b;
    merge_sorted(a.leaves, b.leaves, map_a, map_b, out.leaves);
    prev_a.swap(map_a);
    prev_b.swap(map_b);

    for (int li = nl - 1; li >= 0; --li) {
        const Level& la = a.levels[li];
        const Level& lb = b.levels[li];
        Level& lo = out.levels[li];
        merge_sorted(la.hashes, lb.hashes, map_a, map_b, lo.hashes);
        const uint64_t q = lo.hashes.size();
        lo.masks.assign(q, 0);
        lo.children.assign(8 * q, kEmpty);
        std::vector<char> written(q, 0);
        auto scatter = [&](const Level& src, const std::vector<uint32_t>& map, const std::vector<uint32_t>& child_map) {
            for (uint64_t i = 0; i < src.hashes.size(); ++i) {
                const uint32_t r = map[i];
                uint32_t ch[8];
                for (int k = 0; k < 8; ++k) {
                    const uint32_t c = src.children[8 * i + k];
                    ch[k] = c == kEmpty ? kEmpty : child_map[c];
                }
                if (written[r] && (lo.masks[r] != src.masks[i] || std::memcmp(&lo.children[8 * r], ch, 32) != 0))
                    ++out.collisions;
                lo.masks[r] = src.masks[i];
                std::memcpy(&lo.children[8 * r], ch, 32);
                written[r] = 1;
            }
        };
        scatter(la, map_a, prev_a);  // ScatterWithTransform A then B (:1284-1285)
        scatter(lb, map_b, prev_b);
        prev_a.swap(map_a);
        prev_b.swap(map_b);
    }
    return out;
}

int merge_dags(std::vector<Dag> dags, int threads, Dag& result)
{
    if (dags.empty()) return fail("no DAG to merge");
    uint64_t c = dags.size();
    while (c > 1 && c % 8 == 0) c /= 8;
    if (c != 1) return fail("DAG count %zu is not a power of 8", dags.size());
    const size_t nl = dags[0].levels.size();
    for (auto& d : dags)
        if (d.levels.size() != nl) return fail("DAGs of different depths");

    while (dags.size() > 1) {
        std::vector<Dag> next(dags.size() / 8);
        int error = 0;
        parallel_for(std::max(1, threads), next.size(), [&](uint64_t gb, uint64_t ge) {
            for (uint64_t g = gb; g < ge; ++g) {
                uint64_t root_hashes[8];
                Dag merged = dags[8 * g];
                for (int k = 0; k < 8; ++k) {
                    const Dag& d = dags[8 * g + k];
                    const auto& top = d.levels[0].hashes;
                    if (top.size() > 1) { error = 1; return; }  // :1476
                    root_hashes[k] = top.empty() ? 0 : top[0];
                    if (k > 0) merged = merge_pair(merged, d);
                }
                // New top node (DAGCompression.cu:1505-1533).
                uint8_t mask = 0;
                uint32_t children[8];
                const auto& mh = merged.levels[0].hashes;
                for (int k = 0; k < 8; ++k) {
                    if (root_hashes[k] == 0) { children[k] = kEmpty; continue; }
                    mask |= uint8_t(1u << k);
                    const auto it = std::lower_bound(mh.begin(), mh.end(), root_hashes[k]);
                    children[k] = uint32_t(it - mh.begin());
                }
                Level top;
                if (mask != 0) {
                    top.masks = {mask};
                    top.children.assign(children, children + 8);
                    top.hashes = {hash_children(root_hashes)};
                }
                merged.levels.insert(merged.levels.begin(), std::move(top));
                next[g] = std::move(merged);
            }
        });
        if (error) return fail("a sub-DAG root level holds more than one node");
        dags.swap(next);
    }
    result = std::move(dags[0]);
    result.unique_keys = result.leaf_positions = 0;
    std::memset(result.pre_dedup, 0, sizeof(result.pre_dedup));
    return 0;
}

inline uint32_t popc8(uint8_t m) { return uint32_t(__builtin_popcount(m)); }

// CreateFinalDAG (DAGCompression.cu:1550-1610) with ComputeChildPositions and
// WriteTo (:834-861) and Utils::TotalSize / ChildOffset (Utils.h:68-75).
int create_final_dag(const Dag& dag, std::vector<uint32_t>& words)
{
    const size_t nl = dag.levels.size();
    std::vector<std::vector<uint32_t>> pos(nl);
    std::vector<uint64_t> level_words(nl, 0);
    uint64_t total = 0;
    for (size_t li = 0; li < nl; ++li) {
        const Level& l = dag.levels[li];
        pos[li].resize(l.masks.size());
        uint64_t acc = 0;
        for (size_t i = 0; i < l.masks.size(); ++i) {
            pos[li][i] = uint32_t(acc);
            acc += 1 + popc8(l.masks[i]);
        }
        level_words[li] = acc;
        total += acc;
    }
    total += 2 * dag.leaves.size();
    if (total > 0xFFFFFFFFull) return fail("final DAG of %llu words does not fit uint32 offsets", (unsigned long long)total);
    words.assign(total, 0xFFFFFFFFu);  // Dag.MemSet(0xFF), :1575
    uint64_t offset = 0;
    for (size_t li = 0; li < nl; ++li) {
        const Level& l = dag.levels[li];
        const uint64_t next_start = offset + level_words[li];
        const bool last = li + 1 == nl;
        for (size_t i = 0; i < l.masks.size(); ++i) {
            const uint64_t base = offset + pos[li][i];
            const uint8_t mask = l.masks[i];
            words[base] = mask;
            for (uint32_t s = 0; s < 8; ++s) {
                const uint32_t c = l.children[8 * i + s];
                if (c == kEmpty) continue;
                const uint32_t child_offset = popc8(uint8_t(mask & ((1u << s) - 1u))) + 1;
                words[base + child_offset] = uint32_t(last ? next_start + 2ull * c : next_start + pos[li + 1][c]);
            }
        }
        offset = next_start;
    }
    std::memcpy(words.data() + offset, dag.leaves.data(), dag.leaves.size() * 8);  // :1597
    return 0;
}

template <class T>
T* dup_array(const std::vector<T>& v)
{
    T* p = static_cast<T*>(std::malloc(std::max<size_t>(1, v.size() * sizeof(T))));
    if (!v.empty()) std::memcpy(p, v.data(), v.size() * sizeof(T));
    return p;
}

void export_dag(const Dag& d, ora_dag* out)
{
    std::memset(out, 0, sizeof(*out));
    out->num_levels = int32_t(d.levels.size());
    out->levels = static_cast<ora_level*>(std::calloc(std::max<size_t>(1, d.levels.size()), sizeof(ora_level)));
    for (size_t i = 0; i < d.levels.size(); ++i) {
        out->levels[i].masks = dup_array(d.levels[i].masks);
        out->levels[i].children = dup_array(d.levels[i].children);
        out->levels[i].hashes = dup_array(d.levels[i].hashes);
        out->levels[i].num_nodes = d.levels[i].hashes.size();
    }
    out->leaves = dup_array(d.leaves);
    out->num_leaves = d.leaves.size();
    out->hash_collisions = d.collisions;
    out->unique_keys = d.unique_keys;
    out->leaf_positions = d.leaf_positions;
    std::memcpy(out->pre_dedup, d.pre_dedup, sizeof(d.pre_dedup));
}

Dag import_dag(const ora_dag* in)
{
    Dag d;
    d.levels.resize(in->num_levels);
    for (int i = 0; i < in->num_levels; ++i) {
        const ora_level& l = in->levels[i];
        d.levels[i].masks.assign(l.masks, l.masks + l.num_nodes);
        d.levels[i].children.assign(l.children, l.children + 8 * l.num_nodes);
        d.levels[i].hashes.assign(l.hashes, l.hashes + l.num_nodes);
    }
    d.leaves.assign(in->leaves, in->leaves + in->num_leaves);
    d.collisions = in->hash_collisions;
    return d;
}

}  // namespace

extern "C" {

void ora_murmur128xN(const uint64_t* data, uint32_t size, uint32_t seed, uint64_t len_xor, uint64_t out[2])
{
    murmur128(data, size, seed, len_xor, out);
}

uint64_t ora_hash_children(const uint64_t child_hashes[8]) { return hash_children(child_hashes); }

int ora_create_dag(const uint64_t* keys, uint64_t n, int32_t depth, int32_t threads, ora_dag* out)
{
    if (!out || (n && !keys)) return fail("null argument");
    Dag d;
    const int rc = create_dag(keys, n, depth, threads, d);
    if (rc) return rc;
    export_dag(d, out);
    return 0;
}

int ora_create_dag_u32(const uint32_t* keys, uint64_t n, int32_t depth, int32_t threads, ora_dag* out)
{
    if (depth > 10) return fail("depth %d needs more than 32 key bits", depth);
    if (!out || (n && !keys)) return fail("null argument");
    std::vector<uint64_t> wide(keys, keys + n);
    return ora_create_dag(wide.data(), n, depth, threads, out);
}

int ora_merge_dags(const ora_dag* dags, uint64_t count, int32_t threads, ora_dag* out)
{
    if (!dags || !out) return fail("null argument");
    std::vector<Dag> in;
    in.reserve(count);
    for (uint64_t i = 0; i < count; ++i) in.push_back(import_dag(&dags[i]));
    Dag d;
    const int rc = merge_dags(std::move(in), threads, d);
    if (rc) return rc;
    export_dag(d, out);
    return 0;
}

int ora_create_final_dag(const ora_dag* dag, uint32_t** words, uint64_t* n_words)
{
    if (!dag || !words || !n_words) return fail("null argument");
    std::vector<uint32_t> w;
    const int rc = create_final_dag(import_dag(dag), w);
    if (rc) return rc;
    *words = dup_array(w);
    *n_words = w.size();
    return 0;
}

void ora_free_dag(ora_dag* dag)
{
    if (!dag) return;
    for (int i = 0; i < dag->num_levels && dag->levels; ++i) {
        std::free(dag->levels[i].masks);
        std::free(dag->levels[i].children);
        std::free(dag->levels[i].hashes);
    }
    std::free(dag->levels);
    std::free(dag->leaves);
    std::memset(dag, 0, sizeof(*dag));
}

void ora_free(void* p) { std::free(p); }

const char* ora_last_error(void) { return g_error.c_str(); }

}  // extern "C"
