"""Served mode (dag_compression_fast_amd.serve): requests admitted by
continuous batching into rounds over a resident global DAG.

Integer output, tolerance 0:
- each request's result (its chunk's root hash in the global DAG) equals the
  root hash of the chunk built alone, and the chunk's voxels in the served
  global DAG equal the chunk's own voxels (alone == in the mix);
- the served global DAG equals the oracle's build of the union of the chunks
  (MergeDAGs of the chunks, DAGCompression.cu:1447-1548) and the batched build;
- rounds mix several requests, and the HBM budget splits the work into rounds.
"""
import numpy as np
import pytest

from dagcheck import assert_dag_equal, decode_words


def make_requests(svd, rng, chunk_depth, slot_levels, count, nmax, rate, device_share):
    import test_gpu_parity as T
    from dag_compression_fast_amd import serve
    slots = rng.choice(8 ** slot_levels, size=count, replace=False)
    arrivals = serve.poisson_arrivals(rate, count, seed=5)
    chunks, reqs, bufs = {}, [], []
    for j, s in enumerate(slots):
        keys = T.clustered_keys(rng, chunk_depth, int(rng.integers(1, nmax)))
        chunks[int(s)] = keys
        if j % device_share == 0:
            buf = svd.DeviceBuffer.from_numpy(keys)
            bufs.append(buf)
            reqs.append(serve.Request(j, int(s), buf, float(arrivals[j])))
        else:
            reqs.append(serve.Request(j, int(s), keys, float(arrivals[j])))
    return chunks, reqs, bufs


def serve_and_check(svd, oracle, chunk_depth, slot_levels, count, nmax, rate, budget_keys):
    from dag_compression_fast_amd import serve
    rng = np.random.default_rng(90 + chunk_depth)
    chunks, reqs, bufs = make_requests(svd, rng, chunk_depth, slot_levels, count, nmax, rate, device_share=2)
    total = sum(len(c) for c in chunks.values())
    depth = chunk_depth + slot_levels
    probe = serve.ChunkServer(chunk_depth, slot_levels, max_nodes=total)
    budget = probe.footprint(budget_keys)
    probe.close()
    server = serve.ChunkServer(chunk_depth, slot_levels, max_nodes=total, budget_bytes=budget)
    try:
        results = server.run(reqs)
        assert [r.index for r in results] == list(range(len(reqs)))
        assert all(r.latency >= 0 and r.done >= r.start for r in results)
        rounds = {}
        for r in results:
            rounds.setdefault(r.round, []).append(r)
        assert len(rounds) > 1, "the budget splits the requests into rounds"
        assert max(len(v) for v in rounds.values()) > 1, "some round mixes several requests"
        for rr in rounds.values():
            assert server.footprint(sum(r.n for r in rr)) <= budget or len(rr) == 1
        # Alone == in the mix: root hash of each chunk built alone.
        for r in results:
            alone = svd.create_dag(chunks[r.slot], chunk_depth).to_numpy()
            assert r.root_hash == int(alone["levels"][0]["hashes"][0]), f"slot {r.slot}"
        # The global DAG: the oracle's build of the union, and the batched build.
        got = server.dag()
        words = svd.create_final_dag(got)
        union = np.concatenate([(np.uint64(s) << np.uint64(3 * chunk_depth)) | c for s, c in sorted(chunks.items())])
        exp = oracle.create_dag(union, depth, threads=8)
        assert_dag_equal(got.to_numpy(), exp, "served global DAG == global build")
        assert np.array_equal(words, oracle.create_final_dag(exp))
        per_slot = [chunks.get(s, np.zeros(0, dtype=np.uint64)) for s in range(8 ** slot_levels)]
        offsets = np.zeros(len(per_slot) + 1, dtype=np.uint64)
        offsets[1:] = np.cumsum([len(c) for c in per_slot])
        batched = svd.create_dag_batched(np.concatenate(per_slot), offsets, chunk_depth)
        assert np.array_equal(svd.create_final_dag(batched), words)
        batched.free()
        got.free()
        if depth <= 9:
            # Each chunk's voxels in the served DAG are the chunk's own voxels.
            voxels = np.array(sorted(decode_words(words, depth)), dtype=np.uint64)
            shift = np.uint64(3 * chunk_depth)
            for s, c in chunks.items():
                mine = voxels[(voxels >> shift) == np.uint64(s)] & np.uint64((1 << (3 * chunk_depth)) - 1)
                assert np.array_equal(mine, np.unique(c)), f"voxels of slot {s}"
        stats = serve.latency_stats(results)
        assert stats["p50_ms"] <= stats["p99_ms"] <= stats["max_ms"]
    finally:
        server.close()
        for b in bufs:
            b.free()


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
    serve_and_check(emu, oracle, chunk_depth=5, slot_levels=1, count=7, nmax=600, rate=1e6, budget_keys=1500)


@pytest.mark.gpu
def test_served_chunks_alone_equal_in_the_mix(gpu, oracle):
    print("[serve] 40 chunks d6 in 64 slots", flush=True)
    serve_and_check(gpu, oracle, chunk_depth=6, slot_levels=2, count=40, nmax=30000, rate=20000.0,
                    budget_keys=120000)


@pytest.mark.gpu
def test_served_chunks_depth9_slots512(gpu, oracle):
    print("[serve] 60 chunks d9 in 512 slots", flush=True)
    serve_and_check(gpu, oracle, chunk_depth=9, slot_levels=3, count=60, nmax=200000, rate=5000.0,
                    budget_keys=2_000_000)
