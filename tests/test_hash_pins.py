"""Pins the node hash (MurmurHash128xN of src/Utils.h:119-154) independently of
the oracle's own code.

Chain of evidence:
1. ``dagcheck.murmur3_x86_32`` equals sklearn's MurmurHash3_x86_32 (a library
   implementation) and reproduces SMHasher's published verification value
   0xB0F57EE3 -- this pins the verification harness itself.
2. ``dagcheck.murmur3_x64_128`` reproduces SMHasher's published verification
   value for MurmurHash3_x64_128, 0x6384BA69 -- this pins the block function,
   the finalisation and fmix64 of the 128-bit hash.
3. The oracle's MurmurHash128xN with the canonical length XOR equals the
   canonical hash on 64-byte inputs; the reference variant differs only in
   XOR-ing the block count (4) instead of the byte length (64).
4. The oracle's HashChildrenHashes equals the pinned block function with XOR 4
   (0 mapped to 1), and a planted mistake in the XOR is caught.
"""
import numpy as np
import pytest

import dagcheck

pytestmark = pytest.mark.filterwarnings("ignore")


def test_x86_32_matches_sklearn():
    sk = pytest.importorskip("sklearn.utils")
    rng = np.random.default_rng(3)
    for n in [0, 1, 2, 3, 4, 5, 7, 8, 13, 64, 100]:
        data = bytes(rng.integers(0, 256, size=n, dtype=np.uint8))
        seed = int(rng.integers(0, 2 ** 31))
        ours = dagcheck.murmur3_x86_32(data, seed)
        theirs = sk.murmurhash3_32(data, seed=seed, positive=True)
        assert ours == theirs


def test_smhasher_verification_x86_32():
    v = dagcheck.smhasher_verification(lambda d, s: dagcheck.murmur3_x86_32(d, s).to_bytes(4, "little"), 4)
    assert v == 0xB0F57EE3


def test_smhasher_verification_x64_128():
    def h(d, s):
        h1, h2 = dagcheck.murmur3_x64_128(d, s)
        return h1.to_bytes(8, "little") + h2.to_bytes(8, "little")
    assert dagcheck.smhasher_verification(h, 16) == 0x6384BA69


def test_oracle_block_function_is_canonical(oracle):
    rng = np.random.default_rng(11)
    for _ in range(200):
        blocks = rng.integers(0, 2 ** 64, size=8, dtype=np.uint64)
        seed = int(rng.integers(0, 2 ** 32))
        got = oracle.murmur128xN(blocks, seed, 64)  # canonical: XOR the byte length
        assert got == dagcheck.murmur3_x64_128(blocks.tobytes(), seed)


def test_reference_variant_xors_the_block_count(oracle):
    rng = np.random.default_rng(12)
    for _ in range(200):
        ch = rng.integers(0, 2 ** 64, size=8, dtype=np.uint64)
        ch[rng.random(8) < 0.4] = 0  # empty slots hash as 0
        expect = dagcheck.node_hash(ch)
        assert oracle.hash_children(ch) == expect
        assert oracle.murmur128xN(ch, 0, 4)[0] == expect  # nonzero in practice
        # Canonical MurmurHash3 (XOR 64) gives a different value: the quirk matters.
        assert oracle.murmur128xN(ch, 0, 64)[0] != expect


def test_planted_xor_mistake_is_detected(oracle):
    ch = np.arange(1, 9, dtype=np.uint64)
    good = dagcheck.node_hash(ch)
    assert oracle.murmur128xN(ch, 0, 5)[0] != good
    assert oracle.murmur128xN(ch, 1, 4)[0] != good  # seed matters too
