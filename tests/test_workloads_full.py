"""Full-size workloads on the GPU against the oracle, bit for bit.

- chunks512: the 512-chunk batched build (169M keys, chunk depth 9 -> depth
  12) equals the oracle's CreateDAG of the union of the chunks (MergeDAGs of
  the per-chunk DAGs equals the build of the union, DAGCompression.cu:1447-1548);
- noise1024: the low-redundancy value-noise grid (5.4e8 voxels, depth 10),
  built as 8 batches of 512^3 and merged on the device, equals the oracle's
  per-batch builds merged by the oracle's MergeDAGs;
- noise512 with 32-bit keys equals the oracle's build of the same keys.
"""
import numpy as np
import pytest

from dagcheck import assert_dag_equal

pytestmark = pytest.mark.gpu

THREADS = 16


def test_chunks512_batched_matches_oracle_union(gpu, oracle):
    from dag_compression_fast_amd import workloads as W
    chunks = W.chunk_workload(THREADS)
    offsets = W.chunk_offsets(chunks)
    print(f"[workloads] chunks512 batched: {int(offsets[-1])} keys", flush=True)
    dag = gpu.create_dag_batched(np.concatenate(chunks), offsets, W.CHUNKS["depth"])
    got = dag.to_numpy()
    words = gpu.create_final_dag(dag)
    dag.free()
    union = W.chunk_union(chunks, W.CHUNKS["depth"])
    del chunks
    print("[workloads] chunks512 oracle of the union at depth 12", flush=True)
    exp = oracle.create_dag(union, W.CHUNKS["depth"] + W.CHUNKS["levels_above"], threads=THREADS)
    assert_dag_equal(got, exp, "chunks512 batched")
    assert np.array_equal(words, oracle.create_final_dag(exp))


def test_noise1024_batches_merged_match_oracle(gpu, oracle):
    from dag_compression_fast_amd import workloads as W
    batches = W.noise_batches(THREADS)
    sub = W.NOISE["batch_depth"]
    print(f"[workloads] noise1024: {sum(len(b) for b in batches)} keys in {len(batches)} batches", flush=True)
    dags = []
    for i, b in enumerate(batches):
        d = gpu.create_dag(b, sub)
        exp_b = oracle.create_dag(b, sub, threads=THREADS)
        assert_dag_equal(d.to_numpy(), exp_b, f"noise batch {i}")
        assert d.pre_dedup == exp_b["pre_dedup"]
        dags.append((d, exp_b))
    merged = gpu.merge_dags([d for d, _ in dags])
    exp = oracle.merge_dags([e for _, e in dags], threads=THREADS)
    assert_dag_equal(merged.to_numpy(), exp, "noise1024 merged")
    assert np.array_equal(gpu.create_final_dag(merged), oracle.create_final_dag(exp))
    # The batched build of the same batches is the same DAG.
    offsets = np.zeros(len(batches) + 1, dtype=np.uint64)
    offsets[1:] = np.cumsum([len(b) for b in batches])
    batched = gpu.create_dag_batched(np.concatenate(batches), offsets, sub)
    assert_dag_equal(batched.to_numpy(), exp, "noise1024 batched")


def test_noise512_u32_matches_oracle(gpu, oracle):
    from dag_compression_fast_amd import workloads as W
    keys = W.noise_u32(THREADS)
    print(f"[workloads] noise512 u32: {len(keys)} keys", flush=True)
    exp = oracle.create_dag(keys, W.NOISE_U32["depth"], threads=THREADS)
    dag = gpu.create_dag(keys, W.NOISE_U32["depth"])
    assert_dag_equal(dag.to_numpy(), exp, "noise512 u32")
    assert dag.pre_dedup == exp["pre_dedup"]
    assert np.array_equal(gpu.create_final_dag(dag), oracle.create_final_dag(exp))
    # The same keys as uint64 give the same DAG.
    assert_dag_equal(gpu.create_dag(keys.astype(np.uint64), W.NOISE_U32["depth"]).to_numpy(), exp, "noise512 u64")
