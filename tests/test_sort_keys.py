"""svdag_sort_keys, the counterpart of the reference's SortKeys
(src/DAGCompression.cu:336-357): keys sorted on their low `bits` bits.  Pinned
by numpy's sort (keys below 2^bits must come out exactly as np.sort gives them;
wider keys must come out as a permutation ordered on the low bits), on sizes
around the tile boundaries, on digit-width edge cases and on the error paths."""
import numpy as np
import pytest

CASES = [(1, 9), (100, 9), (4095, 18), (4096, 18), (4097, 33), (70001, 33), (50000, 64), (30000, 1)]


def check_sort(m, rng, n, bits):
    hi = 1 << bits if bits < 64 else None
    keys = (rng.integers(0, hi, size=n, dtype=np.uint64) if hi
            else rng.integers(0, 2 ** 63, size=n, dtype=np.uint64) * np.uint64(2) + np.uint64(1))
    keys[: n // 4] = keys[n // 2: n // 2 + n // 4]  # duplicates
    got = m.sort_keys(keys, bits)
    assert np.array_equal(got, np.sort(keys)), f"n={n} bits={bits}"
    if bits < 64:
        # Keys with bits at or above `bits`: ordered on the low bits only.
        wide = keys | (rng.integers(0, 16, size=n, dtype=np.uint64) << np.uint64(60 if bits <= 60 else 63))
        got = m.sort_keys(wide, bits)
        low = got & np.uint64((1 << bits) - 1)
        assert np.all(low[1:] >= low[:-1])
        assert np.array_equal(np.sort(got), np.sort(wide))


def check_errors(m):
    keys = np.arange(10, dtype=np.uint64)
    with pytest.raises(m.SvdagError):
        m.sort_keys(keys, 0)
    with pytest.raises(m.SvdagError):
        m.sort_keys(keys, 65)
    with pytest.raises(m.SvdagError):
        m.sort_keys(keys.astype(np.uint32), 9)
    assert m.sort_keys(np.zeros(0, dtype=np.uint64), 9).size == 0


def test_sort_keys_on_the_cpu_emulator():
    import emu_ctypes
    m = emu_ctypes.use_emulator()
    try:
        rng = np.random.default_rng(31)
        for n, bits in CASES:
            check_sort(m, rng, n, bits)
        check_errors(m)
    finally:
        from dag_compression_fast_amd import _lib
        import dag_compression_fast_amd as svd
        _lib._lib = None
        _lib._hip = None
        svd._default = None


@pytest.mark.gpu
def test_sort_keys_on_gpu(gpu):
    rng = np.random.default_rng(32)
    for n, bits in CASES + [(3_000_000, 33), (2_000_000, 64), (1_500_001, 27)]:
        print(f"[sort_keys] n={n} bits={bits}", flush=True)
        check_sort(gpu, rng, n, bits)
    check_errors(gpu)
