"""Top-digit-first key sort of plain 64-bit builds (SVDAG_OPT_TOP_DIGIT_FIRST).

A pass on the top digit partitions the keys into Morton-ordered buckets and
writes 32-bit bucket-local keys; a segmented sort inside each bucket finishes
the bits the leaf reduce needs sorted.  The DAG must not depend on the sort's
route: with the option forced on (2), forced off (0) and automatic (1), every
build equals the oracle's (SortKeys + the leaf and level stages,
DAGCompression.cu:1024-1092) bit for bit, integer output with tolerance 0.
Covered: every depth where the bits fit (one segmented pass at depths 9-11,
two at 12) and the depth where they do not (13: the LSD sort runs), each top
digit width 4-9, both tile sizes, duplicates, long runs in one bucket, all keys
in one bucket, out-of-range keys, and (CPU emulator) blocks run one at a time
in decreasing order.
"""
import ctypes

import numpy as np
import pytest

from dagcheck import assert_dag_equal

NEVER, AUTO, ALWAYS = 0, 1, 2


def built(m, oracle, keys, depth, mode, radix=None, tile=None, exp=None):
    from dag_compression_fast_amd import _lib
    ctx = m.Context()
    try:
        ctx.set_option(_lib.OPT_TOP_DIGIT_FIRST, mode)
        if radix:
            ctx.set_option(_lib.OPT_RADIX_BITS, radix)
        if tile:
            ctx.set_option(_lib.OPT_SORT_TILE, tile)
        dag = m.create_dag(keys, depth, ctx=ctx)
        got = dag.to_numpy()
        if exp is None:
            exp = oracle.create_dag(keys, depth)
        assert_dag_equal(got, exp, f"depth {depth} mode {mode} radix {radix} tile {tile}")
        assert dag.pre_dedup == exp["pre_dedup"]
        dag.free()
        return exp
    finally:
        ctx.close()


def cases(rng, scale):
    """(keys, depth, what) of the sizes given by scale."""
    out = []
    for depth in (9, 10, 11, 12, 13):
        n = int(scale * 20000)
        keys = rng.integers(0, 1 << (3 * depth), size=n, dtype=np.uint64)
        out.append((np.concatenate([keys, keys[: n // 3]]), depth, f"uniform d{depth}"))
    import test_gpu_parity as T
    out.append((T.clustered_keys(rng, 11, int(scale * 40000)), 11, "clustered d11"))
    # Every key in one top-digit bucket; and long runs of few values in the
    # bucket-local sorted bits.
    n = int(scale * 30000)
    one = (np.uint64(0x155) << np.uint64(24)) | rng.integers(0, 1 << 24, size=n, dtype=np.uint64)
    out.append((one, 11, "one bucket d11"))
    runs = (rng.choice(np.array([3, 3, 0x1FF], dtype=np.uint64), size=n) << np.uint64(24)) | \
           (rng.choice(np.array([7, 8, 500], dtype=np.uint64), size=n) << np.uint64(15)) | \
           rng.integers(0, 1 << 15, size=n, dtype=np.uint64)
    out.append((runs, 11, "runs d11"))
    return out


def run_all(m, oracle, scale, modes=(ALWAYS, NEVER, AUTO), radixes=(4, 5, 6, 7, 8, 9)):
    rng = np.random.default_rng(0x70D1)
    for keys, depth, what in cases(rng, scale):
        print(f"[top-digit] {what}: {len(keys)} keys", flush=True)
        exp = None
        for mode in modes:
            exp = built(m, oracle, keys, depth, mode, exp=exp)
    keys, depth, _ = cases(np.random.default_rng(1), scale)[2]
    exp = None
    for radix in radixes:
        exp = built(m, oracle, keys, depth, ALWAYS, radix=radix, exp=exp)
    built(m, oracle, keys, depth, ALWAYS, tile=8192, exp=exp)


def run_errors(m):
    from dag_compression_fast_amd import _lib
    ctx = m.Context()
    try:
        ctx.set_option(_lib.OPT_TOP_DIGIT_FIRST, ALWAYS)
        keys = np.arange(5000, dtype=np.uint64) * np.uint64(977)
        keys[1234] = np.uint64(1) << np.uint64(33)  # past a depth-11 code
        with pytest.raises(m.SvdagError) as e:
            m.create_dag(keys, 11, ctx=ctx)
        assert e.value.code == 1
        keys[1234] = 5
        m.create_dag(keys, 11, ctx=ctx).free()
    finally:
        ctx.close()


@pytest.fixture(scope="module")
def emu():
    import emu_ctypes
    svd = emu_ctypes.use_emulator()
    yield svd
    from dag_compression_fast_amd import _lib
    _lib._lib = None
    _lib._hip = None
    svd._default = None


def test_top_digit_first_on_the_cpu_emulator(emu, oracle):
    # Small inputs: the automatic choice is the LSD sort, covered elsewhere.
    # Top digits of 6 and 8 bits leave two segmented passes at depth 11.
    run_all(emu, oracle, 0.1, modes=(ALWAYS,), radixes=(6, 8))
    run_errors(emu)


def test_top_digit_partitions_on_the_cpu_emulator(emu, oracle):
    """The top-digit pass splits the tiles into one contiguous run per
    workgroup (the emulator reports 8 workgroups per CU on 2 CUs: 16
    partitions, placed two per XCD).  40k keys leave some partitions empty
    (10 tiles of 4096), 150k keys give every partition two or three tiles."""
    rng = np.random.default_rng(0x9A27)
    for depth, n in ((11, 40_000), (11, 150_000), (10, 150_000)):
        keys = rng.integers(0, 1 << (3 * depth), size=n, dtype=np.uint64)
        keys = np.concatenate([keys, keys[: n // 4]])
        print(f"[top-digit] partitions d{depth}: {len(keys)} keys", flush=True)
        built(emu, oracle, keys, depth, ALWAYS)


def test_top_digit_first_blocks_in_reverse_order(emu, oracle):
    import emu_ctypes
    L = ctypes.CDLL(emu_ctypes.EMU_LIB)
    L.emu_bad_launch_count.restype = ctypes.c_uint64
    L.emu_launch_count.restype = ctypes.c_uint64
    rng = np.random.default_rng(9)
    keys = rng.integers(0, 1 << 27, size=3000, dtype=np.uint64)
    L.emu_set_schedule(1)
    try:
        bad = L.emu_bad_launch_count()
        before = L.emu_launch_count()
        built(emu, oracle, keys, 9, ALWAYS)
        forced = L.emu_launch_count() - before
        before = L.emu_launch_count()
        built(emu, oracle, keys, 9, NEVER)
        lsd = L.emu_launch_count() - before
        assert L.emu_bad_launch_count() == bad
        # The top-digit route launches more sort kernels than the LSD sort
        # (digit starts, tile planning, segmented scans): it did run.
        assert forced > lsd
    finally:
        L.emu_set_schedule(0)


@pytest.mark.gpu
def test_top_digit_first_on_gpu(gpu, oracle):
    run_all(gpu, oracle, 5.0)
    run_errors(gpu)


@pytest.mark.gpu
def test_top_digit_first_large_on_gpu(gpu, oracle):
    """Above the automatic threshold (2^20 keys): auto takes the top-digit
    route; it must equal the LSD route and the oracle."""
    import test_gpu_parity as T
    rng = np.random.default_rng(0x70D2)
    for depth, n in ((10, 3_000_000), (11, 6_000_000), (12, 4_000_000)):
        keys = T.clustered_keys(rng, depth, n)
        print(f"[top-digit] large d{depth}: {len(keys)} keys", flush=True)
        exp = oracle.create_dag(keys, depth, threads=16)
        built(gpu, oracle, keys, depth, AUTO, exp=exp)
        built(gpu, oracle, keys, depth, NEVER, exp=exp)
