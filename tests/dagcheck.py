"""Independent checkers for DAG results (test helpers, no oracle code).

* ``murmur3_x64_128``: canonical MurmurHash3_x64_128 in Python integers (with
  the published tail handling), pinned by the SMHasher verification value.
* ``node_hash``: the reference's HashChildrenHashes (src/DAGCompression.cu:758-763)
  built from the same block function with the final XOR of the block count.
* ``decode_words``: walks the flat layout the way the reference tracer reads it
  (src/DAGTracer.cu:85-176) and returns the set of voxel Morton codes.
"""
import numpy as np

M64 = (1 << 64) - 1
C1 = 0x87C37B91114253D5
C2 = 0x4CF5AD432745937F


def rotl(x, r):
    return ((x << r) | (x >> (64 - r))) & M64


def fmix(k):
    k ^= k >> 33
    k = (k * 0xFF51AFD7ED558CCD) & M64
    k ^= k >> 33
    k = (k * 0xC4CEB9FE1A85EC53) & M64
    k ^= k >> 33
    return k


def _blocks(h1, h2, k1, k2):
    k1 = (k1 * C1) & M64
    k1 = rotl(k1, 31)
    k1 = (k1 * C2) & M64
    h1 ^= k1
    h1 = rotl(h1, 27)
    h1 = (h1 + h2) & M64
    h1 = (h1 * 5 + 0x52DCE729) & M64
    k2 = (k2 * C2) & M64
    k2 = rotl(k2, 33)
    k2 = (k2 * C1) & M64
    h2 ^= k2
    h2 = rotl(h2, 31)
    h2 = (h2 + h1) & M64
    h2 = (h2 * 5 + 0x38495AB5) & M64
    return h1, h2


def _final(h1, h2, length_xor):
    h1 ^= length_xor
    h2 ^= length_xor
    h1 = (h1 + h2) & M64
    h2 = (h2 + h1) & M64
    h1 = fmix(h1)
    h2 = fmix(h2)
    h1 = (h1 + h2) & M64
    h2 = (h2 + h1) & M64
    return h1, h2


def murmur3_x64_128(data: bytes, seed: int):
    """Canonical MurmurHash3_x64_128 (Appleby, SMHasher) -> (h1, h2)."""
    n = len(data)
    h1 = h2 = seed & 0xFFFFFFFF
    nblocks = n // 16
    for i in range(nblocks):
        k1 = int.from_bytes(data[16 * i:16 * i + 8], "little")
        k2 = int.from_bytes(data[16 * i + 8:16 * i + 16], "little")
        h1, h2 = _blocks(h1, h2, k1, k2)
    tail = data[16 * nblocks:]
    k1 = k2 = 0
    t = len(tail)
    if t > 8:
        k2 = int.from_bytes(tail[8:], "little")
        k2 = (k2 * C2) & M64
        k2 = rotl(k2, 33)
        k2 = (k2 * C1) & M64
        h2 ^= k2
    if t > 0:
        k1 = int.from_bytes(tail[:8], "little")
        k1 = (k1 * C1) & M64
        k1 = rotl(k1, 31)
        k1 = (k1 * C2) & M64
        h1 ^= k1
    return _final(h1, h2, n)


def murmur3_x86_32(data: bytes, seed: int):
    """Canonical MurmurHash3_x86_32 (used to validate the verification harness)."""
    M = 0xFFFFFFFF
    c1, c2 = 0xCC9E2D51, 0x1B873593
    h = seed & M
    n = len(data)
    for i in range(n // 4):
        k = int.from_bytes(data[4 * i:4 * i + 4], "little")
        k = (k * c1) & M
        k = ((k << 15) | (k >> 17)) & M
        k = (k * c2) & M
        h ^= k
        h = ((h << 13) | (h >> 19)) & M
        h = (h * 5 + 0xE6546B64) & M
    tail = data[4 * (n // 4):]
    k = int.from_bytes(tail, "little") if tail else 0
    if tail:
        k = (k * c1) & M
        k = ((k << 15) | (k >> 17)) & M
        k = (k * c2) & M
        h ^= k
    h ^= n
    h ^= h >> 16
    h = (h * 0x85EBCA6B) & M
    h ^= h >> 13
    h = (h * 0xC2B2AE35) & M
    h ^= h >> 16
    return h


def smhasher_verification(hash_fn, hashbytes):
    """SMHasher VerificationTest: hash {0}, {0,1}, ... with seed 256-i, hash the
    concatenation with seed 0, return the first 4 bytes little-endian."""
    key = bytearray(256)
    out = bytearray()
    for i in range(256):
        key[i] = i
        out += hash_fn(bytes(key[:i]), 256 - i)
    final = hash_fn(bytes(out), 0)
    return int.from_bytes(final[:4], "little")


def node_hash(child_hashes):
    """HashChildrenHashes: low word of the 4-block Murmur with XOR 4, 0 -> 1."""
    h1 = h2 = 0
    for i in range(4):
        h1, h2 = _blocks(h1, h2, int(child_hashes[2 * i]), int(child_hashes[2 * i + 1]))
    h1, _ = _final(h1, h2, 4)
    return h1 if h1 != 0 else 1


def decode_words(words, depth):
    """Set of voxel codes reachable in the flat layout (DAGTracer.cu:85-176 semantics)."""
    words = np.asarray(words, dtype=np.uint64)
    if len(words) == 0:
        return set()
    out = set()
    stack = [(0, 0, 0)]  # (word index, level, code prefix)
    levels = depth - 2
    while stack:
        idx, level, prefix = stack.pop()
        if level == levels:
            leaf = int(words[idx]) | (int(words[idx + 1]) << 32)
            for b in range(64):
                if leaf >> b & 1:
                    out.add((prefix << 6) | b)
            continue
        mask = int(words[idx])
        assert 0 < mask < 256, f"bad mask {mask} at word {idx}"
        k = 1
        for slot in range(8):
            if mask >> slot & 1:
                stack.append((int(words[idx + k]), level + 1, (prefix << 3) | slot))
                k += 1
    return out


def assert_dag_equal(got, exp, what=""):
    assert len(got["levels"]) == len(exp["levels"]), f"{what}: level count"
    for i, (g, e) in enumerate(zip(got["levels"], exp["levels"])):
        assert len(g["hashes"]) == len(e["hashes"]), f"{what}: level {i} count {len(g['hashes'])} != {len(e['hashes'])}"
        assert np.array_equal(g["hashes"], e["hashes"]), f"{what}: level {i} hashes differ"
        assert np.array_equal(g["masks"], e["masks"]), f"{what}: level {i} masks differ"
        assert np.array_equal(g["children"], e["children"]), f"{what}: level {i} children differ"
    assert np.array_equal(got["leaves"], exp["leaves"]), f"{what}: leaves differ"


def morton(x, y, z, bits=21):
    code = 0
    for b in range(bits):
        code |= ((x >> b) & 1) << (3 * b + 2)
        code |= ((y >> b) & 1) << (3 * b + 1)
        code |= ((z >> b) & 1) << (3 * b)
    return code
