"""The batched build captured as a HIP graph and replayed
(svdag_batched_graph_*).  Every replay must equal the eager batched build of
the same keys bit for bit: several replays in a row with the program's own
eager work and synchronisations between them, then replays after new keys and
offsets (other chunk sizes, within the capacity) are written into the same
buffers, for the per-chunk sort (chunk depth 6) and the global sort (chunk
depth 10).  Offsets past the capacity make that replay's result an argument
error, and the next valid replay is correct again."""
import numpy as np
import pytest

from dagcheck import assert_dag_equal


def chunk_set(rng, depth, count, nmax, empty_every=5):
    import test_gpu_parity as T
    chunks = [T.clustered_keys(rng, depth, int(rng.integers(1, nmax))) if i % empty_every else
              np.zeros(0, dtype=np.uint64) for i in range(count)]
    offsets = np.zeros(count + 1, dtype=np.uint64)
    offsets[1:] = np.cumsum([len(c) for c in chunks])
    return np.concatenate(chunks), offsets


def run_graph(m, oracle, chunk_depth, count, nmax, replays=3):
    from dag_compression_fast_amd import _lib
    rng = np.random.default_rng(600 + chunk_depth)
    keys, offsets = chunk_set(rng, chunk_depth, count, nmax)
    capacity = 2 * len(keys)
    d_keys = m.DeviceBuffer(8 * capacity)
    d_off = m.DeviceBuffer(8 * (count + 1))
    H = _lib.hip()

    def upload(k, o):
        _lib.hip_check(H.hipMemcpy(d_keys.ptr, k.ctypes.data, k.nbytes, _lib.HIP_H2D), "hipMemcpy")
        _lib.hip_check(H.hipMemcpy(d_off.ptr, o.ctypes.data, o.nbytes, _lib.HIP_H2D), "hipMemcpy")

    upload(keys, offsets)
    graph = m.BatchedGraph(d_keys, d_off, count, chunk_depth, capacity)
    try:
        eager = m.create_dag_batched(keys, offsets, chunk_depth)
        want = eager.to_numpy()
        want_words = m.create_final_dag(eager)
        eager.free()
        for r in range(replays):
            print(f"[graph] d{chunk_depth} replay {r}", flush=True)
            graph.launch()
            got = graph.result()
            assert_dag_equal(got.to_numpy(), want, f"replay {r}")
            assert np.array_equal(m.create_final_dag(got), want_words)
            # The program's own eager work between replays.
            other = m.create_dag(keys[: len(keys) // 3] & np.uint64((1 << 18) - 1), 6)
            other.free()
            _lib.hip_check(H.hipDeviceSynchronize(), "hipDeviceSynchronize")
        # New content in the same buffers: other chunk sizes, more keys.
        for r in range(2):
            keys2, offsets2 = chunk_set(rng, chunk_depth, count, nmax // 2 if r == 0 else nmax, empty_every=3 + r)
            while offsets2[-1] > capacity:
                keys2, offsets2 = chunk_set(rng, chunk_depth, count, nmax, empty_every=3 + r)
            upload(keys2, offsets2)
            graph.launch()
            got = graph.result().to_numpy()
            exp = oracle.merge_dags([oracle.create_dag(keys2[int(offsets2[i]):int(offsets2[i + 1])], chunk_depth)
                                     for i in range(count)])
            assert_dag_equal(got, exp, f"new content {r}")
        # Offsets past the capacity: that replay reports an argument error.
        bad = offsets.copy()
        bad[-1] = capacity + 1
        upload(keys, bad)
        graph.launch()
        with pytest.raises(m.SvdagError) as e:
            graph.result()
        assert e.value.code == 1
        upload(keys, offsets)
        graph.launch()
        assert_dag_equal(graph.result().to_numpy(), want, "after a rejected replay")
    finally:
        graph.close()
        d_keys.free()
        d_off.free()


@pytest.fixture(scope="module")
def emu():
    import emu_ctypes
    svd = emu_ctypes.use_emulator()
    yield svd
    from dag_compression_fast_amd import _lib
    _lib._lib = None
    _lib._hip = None
    svd._default = None


def test_graph_replays_on_the_cpu_emulator(emu, oracle):
    run_graph(emu, oracle, 6, 64, 3000, replays=2)
    run_graph(emu, oracle, 10, 8, 6000, replays=1)


@pytest.mark.gpu
@pytest.mark.parametrize("chunk_depth,count,nmax", [(6, 64, 20000), (9, 512, 60000), (10, 8, 200000)])
def test_graph_replays_on_gpu(gpu, oracle, chunk_depth, count, nmax):
    run_graph(gpu, oracle, chunk_depth, count, nmax)
