"""Resident global DAG (served mode's device state): chunks added round by
round to persistent per-level dedup tables.  After every round the resident
DAG must equal the oracle's MergeDAGs (DAGCompression.cu:1447-1548) of the
chunk DAGs added so far (empty slots empty), and each added chunk's root hash
must equal the root hash of that chunk built alone.  Also: chunks with equal
content in different rounds, a chunk at a filled slot, the capacity check, and
the gathered batched build."""
import numpy as np
import pytest

from dagcheck import assert_dag_equal

CHUNK_DEPTH, SLOT_LEVELS = 6, 2
NUM_SLOTS = 8 ** SLOT_LEVELS


def make_chunks(rng):
    import test_gpu_parity as T
    chunks = {}
    for s in rng.choice(NUM_SLOTS, size=40, replace=False):
        chunks[int(s)] = T.clustered_keys(rng, CHUNK_DEPTH, int(rng.integers(1, 6000)))
    # Equal content at two slots (their subtrees dedup across rounds).
    free = [s for s in range(NUM_SLOTS) if s not in chunks]
    chunks[free[0]] = chunks[next(iter(chunks))].copy()
    chunks[free[1]] = np.array([5], dtype=np.uint64)
    return chunks


def round_dag(m, chunks, slots):
    keys = [chunks[s] if s in slots else np.zeros(0, dtype=np.uint64) for s in range(NUM_SLOTS)]
    offsets = np.zeros(NUM_SLOTS + 1, dtype=np.uint64)
    offsets[1:] = np.cumsum([len(k) for k in keys])
    return m.create_dag_batched(np.concatenate(keys), offsets, CHUNK_DEPTH)


def oracle_union(oracle, chunks, slots):
    empty = oracle.create_dag(np.zeros(0, dtype=np.uint64), CHUNK_DEPTH)
    return oracle.merge_dags([oracle.create_dag(chunks[s], CHUNK_DEPTH) if s in slots else empty
                              for s in range(NUM_SLOTS)])


def run_rounds(m, oracle):
    rng = np.random.default_rng(0x5E5)
    chunks = make_chunks(rng)
    order = list(chunks)
    rng.shuffle(order)
    rounds = [order[:1], order[1:9], order[9:10], order[10:30], order[30:]]
    res = m.ResidentDag(CHUNK_DEPTH, SLOT_LEVELS, max_nodes=sum(len(c) for c in chunks.values()))
    try:
        added = set()
        for i, slots in enumerate(rounds):
            print(f"[resident] round {i}: {len(slots)} chunks", flush=True)
            rd = round_dag(m, chunks, set(slots))
            hashes = res.insert(rd)
            rd.free()
            for s in range(NUM_SLOTS):
                if s in slots:
                    alone = oracle.create_dag(chunks[s], CHUNK_DEPTH)
                    assert int(hashes[s]) == int(alone["levels"][0]["hashes"][0]), f"root hash of slot {s}"
                else:
                    assert hashes[s] == 0
            added |= set(slots)
            got = res.dag()
            exp = oracle_union(oracle, chunks, added)
            assert_dag_equal(got.to_numpy(), exp, f"after round {i}")
            assert np.array_equal(m.create_final_dag(got), oracle.create_final_dag(exp))
            got.free()
            nodes, n_chunks = res.counts()
            assert n_chunks == len(added)
            assert int(nodes[-1]) == len(exp["leaves"])
        # A chunk at a filled slot: an argument error, and the object is done.
        rd = round_dag(m, chunks, {rounds[0][0]})
        with pytest.raises(m.SvdagError) as e:
            res.insert(rd)
        assert e.value.code == 1
        with pytest.raises(m.SvdagError):
            res.dag()
        rd.free()
    finally:
        res.close()


def run_capacity(m):
    # Tables sized for 100 nodes: a round with more nodes at a level than that
    # is refused before any device work, and the object stays usable.
    rng = np.random.default_rng(7)
    chunks = {3: rng.integers(0, 8 ** CHUNK_DEPTH, size=5000, dtype=np.uint64), 9: np.array([1], dtype=np.uint64)}
    res = m.ResidentDag(CHUNK_DEPTH, SLOT_LEVELS, max_nodes=100)
    try:
        rd = round_dag(m, chunks, {3})
        with pytest.raises(m.SvdagError) as e:
            res.insert(rd)
        assert e.value.code == 3
        rd.free()
        rd = round_dag(m, chunks, {9})
        res.insert(rd)
        rd.free()
        assert res.counts()[1] == 1
    finally:
        res.close()


def run_gather(m, oracle):
    import test_gpu_parity as T
    rng = np.random.default_rng(11)
    chunks = [T.clustered_keys(rng, CHUNK_DEPTH, int(rng.integers(0, 3000))) if i % 5 else np.zeros(0, np.uint64)
              for i in range(8)]
    bufs = [m.DeviceBuffer.from_numpy(c) if len(c) else None for c in chunks]
    exp = oracle.merge_dags([oracle.create_dag(c, CHUNK_DEPTH) for c in chunks])
    try:
        got = m.create_dag_batched_gather(bufs, CHUNK_DEPTH)
        assert_dag_equal(got.to_numpy(), exp, "gathered batched build")
    finally:
        for b in bufs:
            if b is not None:
                b.free()
    # Chunks already back to back in one buffer (no copy), and out of order
    # in it (copied into chunk order).
    whole = m.DeviceBuffer.from_numpy(np.concatenate(chunks))
    try:
        starts = np.concatenate([[0], np.cumsum([len(c) for c in chunks])])
        views = [m.DeviceView(whole, 8 * int(starts[i]), len(c)) if len(c) else None for i, c in enumerate(chunks)]
        assert_dag_equal(m.create_dag_batched_gather(views, CHUNK_DEPTH).to_numpy(), exp, "contiguous chunks")
        rev = np.concatenate(chunks[::-1])
        rstarts = np.concatenate([[0], np.cumsum([len(c) for c in chunks[::-1]])])
        whole.free()
        whole = m.DeviceBuffer.from_numpy(rev)
        views = [None] * len(chunks)
        for j, c in enumerate(chunks[::-1]):
            if len(c):
                views[len(chunks) - 1 - j] = m.DeviceView(whole, 8 * int(rstarts[j]), len(c))
        assert_dag_equal(m.create_dag_batched_gather(views, CHUNK_DEPTH).to_numpy(), exp, "chunks out of order")
    finally:
        whole.free()


@pytest.fixture(scope="module")
def emu():
    import emu_ctypes
    svd = emu_ctypes.use_emulator()
    yield svd
    from dag_compression_fast_amd import _lib
    _lib._lib = None
    _lib._hip = None
    svd._default = None


def test_resident_rounds_on_the_cpu_emulator(emu, oracle):
    run_rounds(emu, oracle)


def test_resident_capacity_on_the_cpu_emulator(emu):
    run_capacity(emu)


def test_gather_on_the_cpu_emulator(emu, oracle):
    run_gather(emu, oracle)


@pytest.mark.gpu
def test_resident_rounds_on_gpu(gpu, oracle):
    run_rounds(gpu, oracle)


@pytest.mark.gpu
def test_resident_capacity_on_gpu(gpu):
    run_capacity(gpu)


@pytest.mark.gpu
def test_gather_on_gpu(gpu, oracle):
    run_gather(gpu, oracle)
