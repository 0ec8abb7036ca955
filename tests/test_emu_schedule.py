"""No workgroup of the build depends on another workgroup of the same launch.

The CPU kernel emulator runs the product's .hip sources with every block of a
launch executed one at a time in DECREASING block order (emu_set_schedule(1)).
On a GPU, blocks of a launch run in an order the hardware chooses and are not
guaranteed to be resident together; a kernel in which a block waits for, or
reads what is written by, another block of the same launch would stall here
(the emulator's spin guard) or produce a different result.  Every case below
must match the oracle bit for bit under that schedule, and every launch must
have a valid configuration (no empty or oversized grid or block).

The cases are the inputs whose GPU tests were once skipped: the four
single-key closed forms of DAGCompression.cu:1044-1092, 1566-1597, the depth-3
and depth-4 full cubes, the single-key uniform case and every golden vector,
plus merge and batched builds.
"""
import ctypes

import numpy as np
import pytest

from dagcheck import assert_dag_equal

KATS = [
    ([0], 3, [0x1, 0x2, 0x1, 0x0]),
    ([0x1FF], 3, [0x80, 0x2, 0x0, 0x80000000]),
    ([0], 4, [1, 2, 1, 4, 1, 0]),
    ([0xFFF], 4, [0x80, 2, 0x80, 4, 0x0, 0x80000000]),
]


@pytest.fixture(scope="module")
def emu_rev():
    import emu_ctypes
    svd = emu_ctypes.use_emulator()
    L = ctypes.CDLL(emu_ctypes.EMU_LIB)
    L.emu_bad_launch_count.restype = ctypes.c_uint64
    L.emu_launch_count.restype = ctypes.c_uint64
    L.emu_set_schedule(1)
    yield svd, L
    L.emu_set_schedule(0)
    from dag_compression_fast_amd import _lib
    _lib._lib = None
    _lib._hip = None
    svd._default = None


def _compare(svd, L, oracle, keys, depth, what):
    bad = L.emu_bad_launch_count()
    launches = L.emu_launch_count()
    exp = oracle.create_dag(keys, depth)
    dag = svd.create_dag(keys, depth)
    assert_dag_equal(dag.to_numpy(), exp, what)
    assert dag.pre_dedup == exp["pre_dedup"], what
    words = svd.create_final_dag(dag)
    assert np.array_equal(words, oracle.create_final_dag(exp)), what
    assert L.emu_launch_count() > launches
    assert L.emu_bad_launch_count() == bad, f"{what}: invalid launch configuration"
    return words


@pytest.mark.parametrize("keys,depth,words", KATS)
def test_reverse_schedule_closed_forms(emu_rev, oracle, keys, depth, words):
    svd, L = emu_rev
    got = _compare(svd, L, oracle, np.array(keys, dtype=np.uint64), depth, f"KAT {keys} d{depth}")
    assert got.tolist() == words


@pytest.mark.parametrize("depth", [3, 4, 5])
def test_reverse_schedule_full_cube(emu_rev, oracle, depth):
    svd, L = emu_rev
    words = _compare(svd, L, oracle, np.arange(8 ** depth, dtype=np.uint64), depth, f"cube d{depth}")
    assert len(words) == 9 * (depth - 2) + 2


def test_reverse_schedule_single_uniform_key(emu_rev, oracle):
    svd, L = emu_rev
    rng = np.random.default_rng(1003)
    keys = rng.integers(0, 1 << 9, size=1, dtype=np.uint64)
    _compare(svd, L, oracle, keys, 3, "uniform d3 n=1")


def test_reverse_schedule_golden(emu_rev):
    import test_golden
    svd, L = emu_rev
    bad = L.emu_bad_launch_count()
    for name, keys, depth, counts, leaves, words in test_golden.CASES:
        dag = svd.create_dag(keys, depth)
        assert dag.level_counts == counts.tolist(), name
        assert dag.num_leaves == leaves, name
        assert np.array_equal(svd.create_final_dag(dag), words), name
    assert L.emu_bad_launch_count() == bad


def test_reverse_schedule_random_and_multi_pass(emu_rev, oracle):
    """Multi-tile inputs: several sort passes, many leaf-reduce chunks and
    level-build tiles, so cross-tile offsets are exercised in reverse order."""
    svd, L = emu_rev
    rng = np.random.default_rng(11)
    keys = rng.integers(0, 1 << 30, size=20000, dtype=np.uint64)
    _compare(svd, L, oracle, np.concatenate([keys, keys[:5000]]), 10, "uniform d10")
    import test_gpu_parity as T
    _compare(svd, L, oracle, T.clustered_keys(rng, 9, 30000), 9, "clustered d9")


def test_reverse_schedule_merge_and_batched(emu_rev, oracle):
    import test_gpu_parity as T
    svd, L = emu_rev
    bad = L.emu_bad_launch_count()
    T.test_merge_dags_matches_oracle_and_global(svd, oracle)
    T.test_batched_chunk_depths(svd, oracle, 8)
    T.test_batched_chunk_depths(svd, oracle, 10)
    assert L.emu_bad_launch_count() == bad
