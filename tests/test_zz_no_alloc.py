"""svdag_create_dag_noalloc: the build with its scratch in the library's fixed
device array and its result in the caller's buffer (no allocation in the
call), called through the C ABI.  Each call must equal svdag_create_dag's
result bit for bit, on two calls in a row, the second with a smaller input
(the array keeps the first call's contents), and sizes past the array or the
caller's storage must fail with SVDAG_ERR_MEMORY before any device work."""
import ctypes

import numpy as np
import pytest

from dagcheck import assert_dag_equal

ERR_MEMORY = 3


def noalloc_build(m, keys, depth, storage_bytes=None):
    """One svdag_create_dag_noalloc call; returns (rc, Dag or None, storage)."""
    from dag_compression_fast_amd import _lib
    L = _lib.lib()
    ctx = m.default_context()
    n = len(keys)
    d_keys = m.DeviceBuffer.from_numpy(np.ascontiguousarray(keys, dtype=np.uint64))
    need = int(L.svdag_result_bytes(n, depth))
    storage = m.DeviceBuffer(need if storage_bytes is None else storage_bytes)
    raw = _lib.SvdagDag()
    rc = L.svdag_create_dag_noalloc(ctx.handle, d_keys.ptr, n, depth, storage.ptr, storage.nbytes,
                                    ctypes.byref(raw))
    d_keys.free()
    return rc, (m.Dag(raw, ctx) if rc == 0 else None), storage


def check_against_create_dag(m, keys, depth, what):
    print(f"[no-alloc] {what}: depth={depth} n={len(keys)}", flush=True)
    rc, dag, storage = noalloc_build(m, keys, depth)
    assert rc == 0, m._lib.lib().svdag_last_error().decode()
    assert dag.raw.storage is None  # the caller's buffer: nothing for the library to free
    ref = m.create_dag(keys, depth)
    assert_dag_equal(dag.to_numpy(), ref.to_numpy(), what)
    assert dag.pre_dedup == ref.pre_dedup and dag.leaf_positions == ref.leaf_positions
    assert np.array_equal(m.create_final_dag(dag), m.create_final_dag(ref)), what
    dag.free()
    ref.free()
    storage.free()


def run_cases(m):
    import test_gpu_parity as T
    rng = np.random.default_rng(0x0A110C)
    # First a larger input, then a smaller one: the array holds the first
    # call's contents when the second starts.
    check_against_create_dag(m, T.clustered_keys(rng, 9, 300000), 9, "clustered d9, first call")
    check_against_create_dag(m, T.clustered_keys(rng, 8, 20000), 8, "clustered d8, second, smaller call")
    keys = rng.integers(0, 1 << 30, size=5000, dtype=np.uint64)
    check_against_create_dag(m, np.concatenate([keys, keys[:100]]), 10, "uniform d10, third call")
    check_against_create_dag(m, np.array([0x1FF], dtype=np.uint64), 3, "single key")


def run_errors(m):
    from dag_compression_fast_amd import _lib
    L = _lib.lib()
    # Past the fixed scratch: 2M keys at depth 12 need about 630 MiB; the call
    # fails on the host before any device work.
    n = 2_000_000
    d_keys = m.DeviceBuffer.from_numpy(np.zeros(n, dtype=np.uint64))
    storage = m.DeviceBuffer(int(L.svdag_result_bytes(n, 12)))
    raw = _lib.SvdagDag()
    rc = L.svdag_create_dag_noalloc(m.default_context().handle, d_keys.ptr, n, 12, storage.ptr, storage.nbytes,
                                    ctypes.byref(raw))
    d_keys.free()
    storage.free()
    assert rc == ERR_MEMORY and "scratch" in L.svdag_last_error().decode()
    # Storage smaller than svdag_result_bytes.
    keys = np.arange(0, 8 ** 6, 5, dtype=np.uint64)
    rc, dag, storage = noalloc_build(m, keys, 6, storage_bytes=int(L.svdag_result_bytes(len(keys), 6)) - 256)
    storage.free()
    assert rc == ERR_MEMORY and "storage" in L.svdag_last_error().decode()
    # Empty input: no device work, empty levels.
    raw = _lib.SvdagDag()
    assert L.svdag_create_dag_noalloc(m.default_context().handle, None, 0, 8, None, 0, ctypes.byref(raw)) == 0
    assert raw.num_levels == 6 and raw.num_leaves == 0


def test_noalloc_on_the_cpu_emulator():
    import emu_ctypes
    m = emu_ctypes.use_emulator()
    try:
        run_cases(m)
        run_errors(m)
    finally:
        from dag_compression_fast_amd import _lib
        _lib._lib = None
        _lib._hip = None
        m._default = None


@pytest.mark.gpu
def test_noalloc_on_gpu(gpu):
    run_cases(gpu)
    run_errors(gpu)
