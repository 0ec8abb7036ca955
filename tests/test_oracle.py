"""Pins the CPU oracle against facts that do not come from the oracle itself.

* Closed-form outputs (single voxels, full cubes) worked out by hand from
  src/DAGCompression.cu:1044-1092 and :1566-1597.
* Structural dedup: distinct subtrees per level counted with exact content
  tuples (no hashing), compared with the oracle's per-level node counts.
* Layout: decoding the flat words the way the reference tracer reads them
  (src/DAGTracer.cu:85-176) gives back exactly the input voxel set.
* Hash order and hash values: every level strictly ascending, every node hash
  equal to the SMHasher-pinned node hash of its children's hashes.
* Invariances: permutation and duplicate invariance; MergeDAGs of chunk DAGs
  equals the build of the union (src/DAGCompression.cu:1447-1548).
A planted mistake (a wrong child slot in the layout walk) is shown to fail.
"""
import numpy as np
import pytest

import dagcheck

EMPTY = 0xFFFFFFFF


def kat(oracle, keys, depth):
    return oracle.create_final_dag(oracle.create_dag(np.array(keys, dtype=np.uint64), depth)).tolist()


def test_single_voxel_closed_forms(oracle):
    assert kat(oracle, [0], 3) == [0x1, 0x2, 0x1, 0x0]
    assert kat(oracle, [0x1FF], 3) == [0x80, 0x2, 0x0, 0x80000000]
    assert kat(oracle, [0], 4) == [1, 2, 1, 4, 1, 0]
    assert kat(oracle, [0xFFF], 4) == [0x80, 2, 0x80, 4, 0x0, 0x80000000]


@pytest.mark.parametrize("depth", [3, 4, 5, 6])
def test_full_cube_closed_form(oracle, depth):
    w = kat(oracle, np.arange(8 ** depth), depth)
    assert len(w) == 9 * (depth - 2) + 2
    for lvl in range(depth - 2):
        base = 9 * lvl
        assert w[base] == 0xFF
        assert w[base + 1: base + 9] == [base + 9] * 8
    assert w[-2:] == [0xFFFFFFFF, 0xFFFFFFFF]


def structural_counts(keys, depth):
    """Distinct subtrees per level by exact content (hash-free)."""
    vox = set(int(k) for k in keys)
    blocks = {}
    for v in vox:
        blocks[v >> 6] = blocks.get(v >> 6, 0) | (1 << (v & 63))
    ids = {}
    cur = {}
    for code, m in blocks.items():
        cur[code] = ids.setdefault(("leaf", m), len(ids))
    n_leaves = len(set(blocks.values()))
    counts = []
    for _ in range(depth - 2):
        par = {}
        for code, ident in cur.items():
            par.setdefault(code >> 3, [None] * 8)[code & 7] = ident
        cur = {p: ids.setdefault(("node", tuple(ch)), len(ids)) for p, ch in par.items()}
        counts.append(len(set(cur.values())))
    return list(reversed(counts)), n_leaves


def check_structure(oracle, keys, depth):
    d = oracle.create_dag(keys, depth)
    counts, n_leaves = structural_counts(keys, depth)
    assert [len(lv["hashes"]) for lv in d["levels"]] == counts
    assert len(d["leaves"]) == n_leaves
    assert d["hash_collisions"] == 0
    # Leaves strictly ascending; levels strictly ascending by hash; hashes pinned.
    assert np.all(d["leaves"][1:] > d["leaves"][:-1])
    below = d["leaves"]
    for lv in reversed(d["levels"]):
        h = lv["hashes"]
        assert np.all(h[1:] > h[:-1])
        for i in range(len(h)):
            ch = lv["children"][i]
            m = int(lv["masks"][i])
            assert m != 0
            child_hashes = []
            for s in range(8):
                assert (ch[s] == EMPTY) == (not (m >> s) & 1)
                child_hashes.append(0 if ch[s] == EMPTY else int(below[ch[s]]))
            assert int(h[i]) == dagcheck.node_hash(child_hashes)
        below = h
    words = oracle.create_final_dag(d)
    assert dagcheck.decode_words(words, depth) == set(int(k) for k in keys)
    return d, words


@pytest.mark.parametrize("depth,n,seed", [(3, 1, 0), (3, 40, 1), (3, 300, 2), (4, 50, 3), (4, 2000, 4),
                                          (5, 30, 5), (5, 3000, 6), (5, 20000, 7), (6, 5000, 8)])
def test_brute_force_small_depths(oracle, depth, n, seed):
    rng = np.random.default_rng(seed)
    keys = rng.integers(0, 8 ** depth, size=n, dtype=np.uint64)
    check_structure(oracle, keys, depth)


def test_structure_surface_like(oracle):
    keys = dagcheck_sphere(7, 10, 5.0, 25.0)
    check_structure(oracle, keys, 7)


def dagcheck_sphere(depth, count, rmin, rmax, seed=9):
    rng = np.random.default_rng(seed)
    side = 1 << depth
    pts = []
    for _ in range(count):
        c = rng.uniform(0, side, size=3)
        r = rng.uniform(rmin, rmax)
        lo = np.maximum(0, np.floor(c - r - 1)).astype(int)
        hi = np.minimum(side - 1, np.ceil(c + r + 1)).astype(int)
        g = np.stack(np.meshgrid(*[np.arange(a, b + 1) for a, b in zip(lo, hi)], indexing="ij"), -1).reshape(-1, 3)
        d = np.linalg.norm(g + 0.5 - c, axis=1)
        pts.append(g[np.abs(d - r) < 0.5])
    xyz = np.concatenate(pts)
    return np.array([dagcheck.morton(int(x), int(y), int(z)) for x, y, z in xyz], dtype=np.uint64)


def test_permutation_and_duplicate_invariance(oracle):
    rng = np.random.default_rng(21)
    keys = rng.integers(0, 8 ** 7, size=5000, dtype=np.uint64)
    a = oracle.create_dag(keys, 7)
    b = oracle.create_dag(np.concatenate([rng.permutation(keys), keys[:1000], keys[:10]]), 7)
    dagcheck.assert_dag_equal(a, b, "permutation/duplicates")


@pytest.mark.parametrize("k", [1, 2])
def test_merge_equals_global_build(oracle, k):
    rng = np.random.default_rng(30 + k)
    cd = 4
    chunks = []
    for i in range(8 ** k):
        n = 0 if rng.random() < 0.2 else int(rng.integers(1, 200))
        chunks.append(rng.integers(0, 8 ** cd, size=n, dtype=np.uint64))
    merged = oracle.merge_dags([oracle.create_dag(c, cd) for c in chunks])
    union = np.concatenate([(np.uint64(i) << np.uint64(3 * cd)) | c for i, c in enumerate(chunks)])
    dagcheck.assert_dag_equal(merged, oracle.create_dag(union, cd + k), "merge == global")


def test_multithreaded_oracle_matches_single_thread(oracle):
    rng = np.random.default_rng(40)
    keys = rng.integers(0, 8 ** 9, size=300000, dtype=np.uint64)
    dagcheck.assert_dag_equal(oracle.create_dag(keys, 9, threads=1), oracle.create_dag(keys, 9, threads=8), "threads")


def test_oracle_rejects_out_of_range_keys(oracle):
    with pytest.raises(oracle.OracleError):
        oracle.create_dag(np.array([8 ** 5], dtype=np.uint64), 5)


def test_planted_layout_mistake_is_caught(oracle):
    rng = np.random.default_rng(50)
    keys = rng.integers(0, 8 ** 5, size=400, dtype=np.uint64)
    d = oracle.create_dag(keys, 5)
    words = oracle.create_final_dag(d).copy()
    # Swap the first two child pointers of the root: the decoded voxel set changes.
    assert bin(int(words[0])).count("1") >= 2
    words[1], words[2] = words[2], words[1]
    assert dagcheck.decode_words(words, 5) != set(int(k) for k in keys)
