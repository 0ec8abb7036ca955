"""svdag_sort_keys, the counterpart of the reference's SortKeys
(src/DAGCompression.cu:336-357): keys sorted stably on their low `bits` bits,
as cub::DeviceRadixSort::SortKeys does.  Pinned by numpy: keys below 2^bits
must come out exactly as np.sort gives them, and keys with bits at or above
`bits` exactly as a stable argsort of their low bits orders them; on sizes
around the tile boundaries, on digit-width edge cases, for uint64 and uint32
keys, and on the error paths."""
import numpy as np
import pytest

CASES = [(1, 9), (100, 9), (4095, 18), (4096, 18), (4097, 33), (70001, 33), (50000, 64), (30000, 1)]
CASES_U32 = [(1, 9), (4097, 27), (70001, 27), (50000, 32), (30000, 1)]


def check_sort(m, rng, n, bits, dtype=np.uint64, ctx=None):
    width = 8 * np.dtype(dtype).itemsize
    hi = 1 << bits if bits < width else None
    keys = (rng.integers(0, hi, size=n, dtype=np.uint64) if hi
            else rng.integers(0, 2 ** (width - 1), size=n, dtype=np.uint64) * np.uint64(2) + np.uint64(1))
    keys = keys.astype(dtype)
    keys[: n // 4] = keys[n // 2: n // 2 + n // 4]  # duplicates
    got = m.sort_keys(keys, bits, ctx=ctx)
    assert got.dtype == dtype
    assert np.array_equal(got, np.sort(keys)), f"n={n} bits={bits}"
    if bits < width:
        # Keys with bits at or above `bits`: ordered on the low bits, keys equal
        # there in their input order (stable).
        top = min(width - 4, 60) if bits <= width - 4 else width - 1
        wide = keys | (rng.integers(0, 16 if top < width - 1 else 2, size=n, dtype=np.uint64)
                       << np.uint64(top)).astype(dtype)
        got = m.sort_keys(wide, bits, ctx=ctx)
        low = wide & dtype((1 << bits) - 1)
        assert np.array_equal(got, wide[np.argsort(low, kind="stable")]), f"stable n={n} bits={bits}"


def check_errors(m):
    keys = np.arange(10, dtype=np.uint64)
    with pytest.raises(m.SvdagError):
        m.sort_keys(keys, 0)
    with pytest.raises(m.SvdagError):
        m.sort_keys(keys, 65)
    with pytest.raises(m.SvdagError):
        m.sort_keys(keys.astype(np.uint32), 33)
    with pytest.raises(m.SvdagError):
        m.sort_keys(keys.astype(np.int64), 9)
    assert m.sort_keys(np.zeros(0, dtype=np.uint64), 9).size == 0


def check_narrow_digits(m, rng):
    """4-bit digits: after the first pass the runs of one digit are longer than
    a tile, so later passes see tiles uniform in the earlier bits (ranked
    without stability when every key fits the sorted bits, stably otherwise)."""
    from dag_compression_fast_amd import _lib
    ctx = m.Context()
    try:
        ctx.set_option(_lib.OPT_RADIX_BITS, 4)
        check_sort(m, rng, 200000, 24, ctx=ctx)
        check_sort(m, rng, 200000, 20, np.uint32, ctx=ctx)
    finally:
        ctx.close()


def test_sort_keys_on_the_cpu_emulator():
    import emu_ctypes
    m = emu_ctypes.use_emulator()
    try:
        rng = np.random.default_rng(31)
        for n, bits in CASES:
            check_sort(m, rng, n, bits)
        for n, bits in CASES_U32:
            check_sort(m, rng, n, bits, np.uint32)
        check_narrow_digits(m, rng)
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
    for n, bits in CASES_U32 + [(3_000_000, 27), (1_000_000, 32)]:
        print(f"[sort_keys] u32 n={n} bits={bits}", flush=True)
        check_sort(gpu, rng, n, bits, np.uint32)
    check_narrow_digits(gpu, rng)
    check_errors(gpu)
