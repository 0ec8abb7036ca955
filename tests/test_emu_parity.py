"""Kernel logic on the CPU: the product's .hip sources compiled by g++ against a
mock HIP runtime (tests/emu) run through the same C ABI and Python API, and are
compared with the oracle.  Covers the closed forms, small random and clustered
inputs, merge and batched builds, and error paths, without a GPU."""
import numpy as np
import pytest

from dagcheck import assert_dag_equal, decode_words


@pytest.fixture(scope="module")
def emu():
    import emu_ctypes
    svd = emu_ctypes.use_emulator()
    yield svd
    from dag_compression_fast_amd import _lib
    _lib._lib = None
    _lib._hip = None
    svd._default = None


def build_and_compare(emu, oracle, keys, depth):
    exp = oracle.create_dag(keys, depth)
    dag = emu.create_dag(keys, depth)
    assert_dag_equal(dag.to_numpy(), exp)
    assert dag.pre_dedup == exp["pre_dedup"]
    words = emu.create_final_dag(dag)
    assert np.array_equal(words, oracle.create_final_dag(exp))
    return words


@pytest.mark.parametrize("keys,depth,words", [
    ([0], 3, [0x1, 0x2, 0x1, 0x0]),
    ([0x1FF], 3, [0x80, 0x2, 0x0, 0x80000000]),
    ([0], 4, [1, 2, 1, 4, 1, 0]),
    ([0xFFF], 4, [0x80, 2, 0x80, 4, 0x0, 0x80000000]),
])
def test_emu_closed_forms(emu, keys, depth, words):
    dag = emu.create_dag(np.array(keys, dtype=np.uint64), depth)
    assert emu.create_final_dag(dag).tolist() == words


@pytest.mark.parametrize("depth,n", [(3, 1), (3, 66), (5, 3000), (8, 20000), (12, 8000), (21, 3000)])
def test_emu_random(emu, oracle, depth, n):
    rng = np.random.default_rng(depth * 7 + n)
    keys = rng.integers(0, 1 << (3 * depth), size=n, dtype=np.uint64)
    keys = np.concatenate([keys, keys[: n // 3]])
    words = build_and_compare(emu, oracle, keys, depth)
    if depth <= 8:
        assert decode_words(words, depth) == set(int(k) for k in keys)


def test_emu_clustered_and_runs(emu, oracle):
    import test_gpu_parity as T
    rng = np.random.default_rng(3)
    build_and_compare(emu, oracle, T.clustered_keys(rng, 9, 30000), 9)
    build_and_compare(emu, oracle, np.repeat(np.array([7, 64 * 5 + 3, 1 << 20], dtype=np.uint64),
                                             [20000, 9000, 3]), 9)


def test_emu_long_runs_in_sorted_bits(emu, oracle):
    """Later sort passes rank a tile unstably when all its keys agree on the
    bits the earlier passes sorted.  Long runs of few values in those bits make
    most tiles uniform, with mixed tiles at the run boundaries."""
    rng = np.random.default_rng(5)
    n = 60000
    # depth 10: bits 15..30 in two passes (15..23, 23..30)
    d1 = rng.choice(np.array([3, 4, 200], dtype=np.uint64), size=n, p=[0.5, 0.2, 0.3])
    d2 = rng.integers(0, 1 << 7, size=n, dtype=np.uint64)
    low = rng.integers(0, 1 << 15, size=n, dtype=np.uint64)
    keys = (d2 << np.uint64(23)) | (d1 << np.uint64(15)) | low
    build_and_compare(emu, oracle, np.concatenate([keys, keys[:n // 4]]), 10)
    # depth 12: bits 15..36 in three passes of 7; the third sees the first two
    a = rng.choice(np.array([5, 6, 0x3FFF], dtype=np.uint64), size=n, p=[0.4, 0.4, 0.2])
    hi = rng.integers(0, 1 << 7, size=n, dtype=np.uint64)
    keys = (hi << np.uint64(29)) | (a << np.uint64(15)) | rng.integers(0, 1 << 15, size=n, dtype=np.uint64)
    build_and_compare(emu, oracle, keys, 12)


def test_emu_solid_regions(emu, oracle):
    """Long runs of all-ones leaves: the leaf insert probes once per run."""
    import test_gpu_parity as T
    build_and_compare(emu, oracle, T.solid_keys(np.random.default_rng(6), 7, 2000), 7)


def test_emu_u32(emu, oracle):
    rng = np.random.default_rng(4)
    keys = rng.integers(0, 8 ** 9, size=20000, dtype=np.uint64)
    exp = oracle.create_dag(keys, 9)
    assert_dag_equal(emu.create_dag(keys.astype(np.uint32), 9).to_numpy(), exp)


def test_emu_merge_and_batched(emu, oracle):
    import test_gpu_parity as T
    T.test_merge_dags_matches_oracle_and_global(emu, oracle)
    T.test_batched_matches_merge(emu, oracle)
    for chunk_depth in (6, 8, 9, 10):
        T.test_batched_chunk_depths(emu, oracle, chunk_depth)
    for first_bits in (4, 8):
        T.test_batched_segmented_digit_split(emu, oracle, first_bits)
    T.test_batched_many_sparse_chunks(emu, oracle)


def test_emu_errors(emu):
    with pytest.raises(emu.SvdagError) as e:
        emu.create_dag(np.array([1, 2, 1 << 30], dtype=np.uint64), 8)
    assert e.value.code == 1
    with pytest.raises(emu.SvdagError):
        emu.create_dag(np.array([0], dtype=np.uint64), 22)
    dag = emu.create_dag(np.zeros(0, dtype=np.uint64), 8)
    assert dag.level_counts == [0] * 6


def test_emu_host_dag_validation(emu, oracle):
    keys = np.arange(0, 8 ** 5, 7, dtype=np.uint64)
    host = emu.create_dag(keys, 5).to_host()
    # Corrupt a child index in the host copy: upload must refuse it.
    import ctypes
    lv = host.raw.levels[0]
    ch = ctypes.cast(lv.children, ctypes.POINTER(ctypes.c_uint32))
    slot = next(s for s in range(8) if ch[s] != 0xFFFFFFFF)
    ch[slot] = 10 ** 6
    with pytest.raises(emu.SvdagError):
        host.to_device()
