"""Writes tests/golden/svdag_golden.npz: small seeded DAG-build cases with their
expected outputs (per-level node counts, leaf count, flat uint32 layout).

The reference ships no test vectors, so the expected outputs come from the CPU
oracle.  Before a case is written it is checked independently of the oracle:
- the per-level node counts equal the distinct-subtree counts obtained from
  exact content tuples (no hashing);
- decoding the flat layout as the reference tracer reads it
  (src/DAGTracer.cu:85-176) gives back exactly the input voxel set;
- every stored node hash equals the SMHasher-pinned node hash of its
  children's hashes, and each level is strictly ascending.

Run from the repository root: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")]

import dagcheck  # noqa: E402
import oracle_ctypes as oracle  # noqa: E402
from test_oracle import check_structure  # noqa: E402


def clustered(rng, depth, n):
    side = 1 << depth
    pts = []
    while sum(len(p) for p in pts) < n:
        start = rng.integers(0, side, size=3)
        walk = np.clip(start + np.cumsum(rng.integers(-1, 2, size=(int(rng.integers(8, 200)), 3)), axis=0), 0, side - 1)
        pts.append(walk)
    xyz = np.concatenate(pts)[:n]
    keys = np.array([dagcheck.morton(int(x), int(y), int(z)) for x, y, z in xyz], dtype=np.uint64)
    keys = np.concatenate([keys, keys[: n // 4]])
    rng.shuffle(keys)
    return keys


def cases():
    rng = np.random.default_rng(0x5EED)
    yield "kat_d3_origin", np.array([0], dtype=np.uint64), 3
    yield "kat_d3_corner", np.array([0x1FF], dtype=np.uint64), 3
    yield "kat_d4_origin", np.array([0], dtype=np.uint64), 4
    yield "kat_d4_corner", np.array([0xFFF], dtype=np.uint64), 4
    yield "cube_d4", np.arange(8 ** 4, dtype=np.uint64), 4
    yield "uniform_d5", rng.integers(0, 8 ** 5, size=300, dtype=np.uint64), 5
    yield "uniform_d7", rng.integers(0, 8 ** 7, size=3000, dtype=np.uint64), 7
    yield "clustered_d8", clustered(rng, 8, 2500), 8
    yield "uniform_d12", rng.integers(0, 8 ** 12, size=2000, dtype=np.uint64), 12


def main():
    out = {}
    for name, keys, depth in cases():
        d, words = check_structure(oracle, keys, depth)
        out[f"{name}/keys"] = keys
        out[f"{name}/depth"] = np.array([depth], dtype=np.int32)
        out[f"{name}/level_counts"] = np.array([len(lv["hashes"]) for lv in d["levels"]], dtype=np.uint64)
        out[f"{name}/num_leaves"] = np.array([len(d["leaves"])], dtype=np.uint64)
        out[f"{name}/words"] = np.asarray(words, dtype=np.uint32)
        print(f"{name}: depth {depth}, {len(keys)} keys, {len(words)} words")
    path = os.path.join(HERE, "svdag_golden.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
