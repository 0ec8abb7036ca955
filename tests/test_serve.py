"""Served mode (dag_compression_fast_amd.serve): requests admitted continuously
into slots; each request's chunk DAG must equal the chunk built alone, and the
merge of the served chunks must equal the batched build and the oracle's global
build of the same keys (integer output: tolerance 0)."""
import numpy as np
import pytest

from dagcheck import assert_dag_equal


def make_chunks(rng, count, chunk_depth, nmax):
    import test_gpu_parity as T
    chunks = []
    for i in range(count):
        n = int(rng.integers(1, nmax)) if i % 5 else 0  # some sub-volumes stay empty
        chunks.append(T.clustered_keys(rng, chunk_depth, n) if n else np.zeros(0, dtype=np.uint64))
    return chunks


def serve_and_check(svd, oracle, chunks, chunk_depth, slots, rate):
    from dag_compression_fast_amd import serve
    live = [i for i, c in enumerate(chunks) if len(c)]
    arrivals = serve.poisson_arrivals(rate, len(live), seed=5)
    reqs = [serve.Request(j, i, chunks[i], float(arrivals[j])) for j, i in enumerate(live)]
    server = serve.ChunkServer(chunk_depth, slots=slots, max_chunk_keys=max(len(c) for c in chunks))
    try:
        results = server.run(reqs)
        assert [r.index for r in results] == list(range(len(reqs)))
        assert all(r.latency >= 0 for r in results)
        # Each result equals the chunk built alone.
        for r in results:
            alone = svd.create_dag(chunks[r.chunk], chunk_depth)
            assert_dag_equal(r.dag.to_numpy(), alone.to_numpy(), f"chunk {r.chunk} in the mix")
            alone.free()
        merged = server.merge(results, len(chunks))
        got = merged.to_numpy()
        offsets = np.zeros(len(chunks) + 1, dtype=np.uint64)
        offsets[1:] = np.cumsum([len(c) for c in chunks])
        batched = svd.create_dag_batched(np.concatenate(chunks), offsets, chunk_depth).to_numpy()
        assert_dag_equal(got, batched, "served merge == batched")
        k = {8: 1, 64: 2, 512: 3}[len(chunks)]
        union = np.concatenate([(np.uint64(i) << np.uint64(3 * chunk_depth)) | c for i, c in enumerate(chunks)])
        assert_dag_equal(got, oracle.create_dag(union, chunk_depth + k), "served merge == global build")
        merged.free()
        for r in results:
            r.dag.free()
        stats = serve.latency_stats(results)
        assert stats["p50_ms"] <= stats["p99_ms"] <= stats["max_ms"]
    finally:
        server.close()


@pytest.fixture(scope="module")
def emu():
    import emu_ctypes
    svd = emu_ctypes.use_emulator()
    yield svd
    from dag_compression_fast_amd import _lib
    _lib._lib = None
    _lib._hip = None
    svd._default = None


def test_served_chunks_on_the_cpu_emulator(emu, oracle):
    rng = np.random.default_rng(90)
    serve_and_check(emu, oracle, make_chunks(rng, 8, 5, 600), 5, slots=2, rate=5000.0)


@pytest.mark.gpu
def test_served_chunks_equal_alone_and_batched(gpu, oracle):
    rng = np.random.default_rng(91)
    print("[serve] 64 chunks d6, 4 slots", flush=True)
    serve_and_check(gpu, oracle, make_chunks(rng, 64, 6, 30000), 6, slots=4, rate=20000.0)
