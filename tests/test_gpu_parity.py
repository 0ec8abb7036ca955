"""GPU parity: the HIP library (through its C ABI) against the CPU oracle.

Integer work, so every comparison is bit-exact: per-level node counts, hashes,
masks, child indices, leaves, and the flat uint32 layout.
"""
import numpy as np
import pytest

from dagcheck import assert_dag_equal, decode_words

pytestmark = pytest.mark.gpu


def clustered_keys(rng, depth, n, dup=0.3):
    """Surface-like keys: random walks of voxels (dense leaves) plus duplicates."""
    side = 1 << depth
    pts = []
    while sum(len(p) for p in pts) < n:
        start = rng.integers(0, side, size=3)
        steps = rng.integers(-1, 2, size=(int(rng.integers(8, 400)), 3))
        walk = np.clip(start + np.cumsum(steps, axis=0), 0, side - 1)
        pts.append(walk)
    xyz = np.concatenate(pts)[:n].astype(np.uint64)
    code = np.zeros(len(xyz), dtype=np.uint64)
    for b in range(depth):
        for a in range(3):
            code |= ((xyz[:, a] >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b + 2 - a)
    extra = rng.choice(code, size=int(len(code) * dup / (1 - dup))) if len(code) else code
    keys = np.concatenate([code, extra])
    rng.shuffle(keys)
    return keys


def check_case(gpu, oracle, keys, depth, what):
    print(f"[gpu-parity] {what}: depth={depth} n={len(keys)}", flush=True)
    exp = oracle.create_dag(keys, depth, threads=8)
    dag = gpu.create_dag(keys, depth)
    got = dag.to_numpy()
    assert_dag_equal(got, exp, what)
    assert dag.pre_dedup == exp["pre_dedup"], f"{what}: pre-dedup counts"
    assert dag.leaf_positions == exp["leaf_positions"]
    words = gpu.create_final_dag(dag)
    exp_words = oracle.create_final_dag(exp)
    assert np.array_equal(words, exp_words), f"{what}: final layout differs"
    return dag, words


# Inputs that once stalled under an earlier kernel design whose workgroups
# waited on each other (decoupled look-back, tile tickets); that design is gone
# and no kernel waits on another workgroup (tests/test_emu_schedule.py runs
# these inputs with every launch's blocks one at a time in reverse order).
KATS = [
    ([0], 3, [0x1, 0x2, 0x1, 0x0]),
    ([0x1FF], 3, [0x80, 0x2, 0x0, 0x80000000]),
    ([0], 4, [1, 2, 1, 4, 1, 0]),
    ([0xFFF], 4, [0x80, 2, 0x80, 4, 0x0, 0x80000000]),
]


@pytest.mark.parametrize("keys,depth,words", KATS)
def test_closed_form_kats(gpu, keys, depth, words):
    print(f"[gpu-parity] KAT depth={depth} keys={keys}", flush=True)
    dag = gpu.create_dag(np.array(keys, dtype=np.uint64), depth)
    assert gpu.create_final_dag(dag).tolist() == words


@pytest.mark.parametrize("depth", [3, 4, 5, 6])
def test_full_cube(gpu, depth):
    print(f"[gpu-parity] full cube depth={depth}", flush=True)
    keys = np.arange(8 ** depth, dtype=np.uint64)
    dag = gpu.create_dag(keys, depth)
    words = gpu.create_final_dag(dag)
    assert len(words) == 9 * (depth - 2) + 2
    assert words[-2] == 0xFFFFFFFF and words[-1] == 0xFFFFFFFF


@pytest.mark.parametrize("depth,n", [(3, 1), (3, 50), (4, 300), (5, 5000), (6, 20000), (8, 100000),
                                     (10, 200000), (12, 300000), (16, 200000), (21, 100000)])
def test_random_uniform(gpu, oracle, depth, n):
    rng = np.random.default_rng(1000 + depth)
    keys = rng.integers(0, 1 << (3 * depth), size=n, dtype=np.uint64)
    keys = np.concatenate([keys, keys[: n // 3]])
    rng.shuffle(keys)
    check_case(gpu, oracle, keys, depth, f"uniform d{depth}")


@pytest.mark.parametrize("depth,n", [(6, 30000), (8, 200000), (9, 500000), (11, 300000)])
def test_random_clustered(gpu, oracle, depth, n):
    rng = np.random.default_rng(2000 + depth)
    check_case(gpu, oracle, clustered_keys(rng, depth, n), depth, f"clustered d{depth}")


def test_sphere_shells_256(gpu, oracle):
    keys = gpu.gen_sphere_shells(0x5EED0000, 8, 64, 8.0, 40.0, 0.2)
    dag, words = check_case(gpu, oracle, keys, 8, "sphere shells 256^3")
    assert decode_words(words, 8) == set(np.unique(keys).tolist())


def test_u32_keys(gpu, oracle):
    keys = gpu.gen_sphere_shells(0x5EED0002, 9, 40, 16.0, 80.0, 0.25).astype(np.uint32)
    print(f"[gpu-parity] u32 d9 n={len(keys)}", flush=True)
    exp = oracle.create_dag(keys.astype(np.uint64), 9, threads=8)
    dag = gpu.create_dag(keys, 9)
    assert_dag_equal(dag.to_numpy(), exp, "u32")


def test_all_duplicates(gpu, oracle):
    keys = np.full(100000, 123456789, dtype=np.uint64)
    check_case(gpu, oracle, keys, 10, "all duplicates")


def test_long_duplicate_runs_across_tiles(gpu, oracle):
    # Runs of one leaf block far longer than a tile.
    keys = np.repeat(np.array([7, 64 * 5 + 3, 1 << 20], dtype=np.uint64), [20000, 9000, 3])
    check_case(gpu, oracle, keys, 9, "long runs")


def test_empty_input(gpu):
    dag = gpu.create_dag(np.zeros(0, dtype=np.uint64), 8)
    assert dag.num_levels == 6 and dag.level_counts == [0] * 6 and dag.num_leaves == 0


def test_key_out_of_range_is_an_error(gpu):
    keys = np.array([1, 2, 1 << 30], dtype=np.uint64)
    with pytest.raises(gpu.SvdagError) as e:
        gpu.create_dag(keys, 8)
    assert e.value.code == 1


@pytest.mark.parametrize("depth", [2, 22])
def test_bad_depth_is_an_error(gpu, depth):
    with pytest.raises(gpu.SvdagError):
        gpu.create_dag(np.array([0], dtype=np.uint64), depth)


def test_to_host_round_trip(gpu, oracle):
    keys = gpu.gen_sphere_shells(7, 7, 10, 4.0, 20.0, 0.2)
    dag = gpu.create_dag(keys, 7)
    host = dag.to_host()
    assert_dag_equal(host.to_numpy(), dag.to_numpy(), "to_host")
    back = host.to_device()
    assert np.array_equal(gpu.create_final_dag(back), gpu.create_final_dag(dag))


def test_repeat_is_deterministic(gpu):
    rng = np.random.default_rng(5)
    keys = clustered_keys(rng, 10, 200000)
    a = gpu.create_dag(keys, 10).to_numpy()
    b = gpu.create_dag(keys, 10).to_numpy()
    assert_dag_equal(a, b, "repeat")


def test_merge_dags_matches_oracle_and_global(gpu, oracle):
    rng = np.random.default_rng(77)
    chunk_depth, k = 6, 1
    chunks = []
    for i in range(8 ** k):
        n = 0 if i == 3 else int(rng.integers(1, 4000))
        chunks.append(clustered_keys(rng, chunk_depth, n) if n else np.zeros(0, dtype=np.uint64))
    print("[gpu-parity] merge 8 chunks d6", flush=True)
    gdags = [gpu.create_dag(c, chunk_depth) for c in chunks]
    merged = gpu.merge_dags(gdags).to_numpy()
    odags = [oracle.create_dag(c, chunk_depth) for c in chunks]
    exp = oracle.merge_dags(odags)
    assert_dag_equal(merged, exp, "merge")
    union = np.concatenate([(np.uint64(i) << np.uint64(3 * chunk_depth)) | c for i, c in enumerate(chunks)])
    assert_dag_equal(merged, oracle.create_dag(union, chunk_depth + k), "merge == global")


def test_batched_matches_merge(gpu, oracle):
    rng = np.random.default_rng(78)
    chunk_depth, count = 7, 64
    chunks = []
    for i in range(count):
        n = 0 if i % 9 == 4 else int(rng.integers(1, 20000))
        chunks.append(clustered_keys(rng, chunk_depth, n) if n else np.zeros(0, dtype=np.uint64))
    offsets = np.zeros(count + 1, dtype=np.uint64)
    offsets[1:] = np.cumsum([len(c) for c in chunks])
    print("[gpu-parity] batched 64 chunks d7", flush=True)
    batched = gpu.create_dag_batched(np.concatenate(chunks), offsets, chunk_depth)
    exp = oracle.merge_dags([oracle.create_dag(c, chunk_depth) for c in chunks])
    assert_dag_equal(batched.to_numpy(), exp, "batched")
    merged = gpu.merge_dags([gpu.create_dag(c, chunk_depth) for c in chunks])
    assert np.array_equal(gpu.create_final_dag(batched), gpu.create_final_dag(merged))


@pytest.mark.parametrize("chunk_depth", [6, 8, 9, 10])
def test_batched_chunk_depths(gpu, oracle, chunk_depth):
    """Chunk-local sort spans of 3, 9 and 12 bits (sorted inside each chunk in
    one or two passes) and 15 bits (the global sort); chunk sizes around the
    4096-key tile, empty and single-key chunks."""
    rng = np.random.default_rng(90 + chunk_depth)
    sizes = [0, 1, 3000, 4096, 4097, 25000, 9000, 130]
    chunks = [clustered_keys(rng, chunk_depth, n) if n else np.zeros(0, dtype=np.uint64) for n in sizes]
    offsets = np.zeros(len(chunks) + 1, dtype=np.uint64)
    offsets[1:] = np.cumsum([len(c) for c in chunks])
    print(f"[gpu-parity] batched 8 chunks d{chunk_depth}", flush=True)
    batched = gpu.create_dag_batched(np.concatenate(chunks), offsets, chunk_depth)
    exp = oracle.merge_dags([oracle.create_dag(c, chunk_depth) for c in chunks])
    assert_dag_equal(batched.to_numpy(), exp, f"batched d{chunk_depth}")


def test_batched_many_sparse_chunks(gpu, oracle):
    """4096 chunks, few of them filled: the segmented sort's 32-bit
    chunk-local keys get their chunk back from a chunk table too large for the
    leaf reduce to stage (more than 512 chunks), across long runs of empty
    chunks."""
    rng = np.random.default_rng(4096)
    chunk_depth, num = 6, 4096
    filled = set(int(c) for c in rng.choice(num, size=40, replace=False)) | {0, 1, num - 1}
    chunks = [clustered_keys(rng, chunk_depth, int(rng.integers(1, 9000))) if c in filled
              else np.zeros(0, dtype=np.uint64) for c in range(num)]
    offsets = np.zeros(num + 1, dtype=np.uint64)
    offsets[1:] = np.cumsum([len(c) for c in chunks])
    print(f"[gpu-parity] batched {num} chunks d{chunk_depth}", flush=True)
    batched = gpu.create_dag_batched(np.concatenate(chunks), offsets, chunk_depth)
    union = np.concatenate([(np.uint64(i) << np.uint64(3 * chunk_depth)) | c for i, c in enumerate(chunks)])
    exp = oracle.create_dag(union, chunk_depth + 4)
    assert_dag_equal(batched.to_numpy(), exp, "batched 4096 sparse chunks")


@pytest.mark.parametrize("bits,tile", [(4, 4096), (6, 8192), (7, 4096), (9, 4096), (9, 8192)])
def test_radix_digit_widths(gpu, oracle, bits, tile):
    from dag_compression_fast_amd import _lib
    ctx = gpu.Context()
    try:
        ctx.set_option(_lib.OPT_RADIX_BITS, bits)
        ctx.set_option(_lib.OPT_SORT_TILE, tile)
        ctx.set_option(_lib.OPT_WAIT_SECONDS, 60)
        rng = np.random.default_rng(3000 + bits + tile)
        keys = clustered_keys(rng, 10, 400000)
        print(f"[gpu-parity] radix bits {bits} tile {tile}: depth=10 n={len(keys)}", flush=True)
        exp = oracle.create_dag(keys, 10, threads=8)
        dag = gpu.create_dag(keys, 10, ctx=ctx)
        assert_dag_equal(dag.to_numpy(), exp, f"radix {bits} tile {tile}")
        assert np.array_equal(gpu.create_final_dag(dag), oracle.create_final_dag(exp))
        dag.free()
    finally:
        ctx.close()


@pytest.mark.parametrize("first_bits", [4, 5, 8])
def test_batched_segmented_digit_split(gpu, oracle, first_bits):
    """The per-chunk sort of chunk depth 9 (12 bits) with another digit split
    (4 + 8, 5 + 7, 8 + 4 bits) gives the same DAG."""
    from dag_compression_fast_amd import _lib
    rng = np.random.default_rng(400 + first_bits)
    sizes = [0, 1, 5000, 4096, 70000, 0, 12000, 300]
    chunks = [clustered_keys(rng, 9, n) if n else np.zeros(0, dtype=np.uint64) for n in sizes]
    offsets = np.zeros(len(chunks) + 1, dtype=np.uint64)
    offsets[1:] = np.cumsum([len(c) for c in chunks])
    ctx = gpu.Context()
    try:
        ctx.set_option(_lib.OPT_SEG_FIRST_BITS, first_bits)
        print(f"[gpu-parity] batched d9 first bits {first_bits}", flush=True)
        batched = gpu.create_dag_batched(np.concatenate(chunks), offsets, 9, ctx=ctx)
        exp = oracle.merge_dags([oracle.create_dag(c, 9) for c in chunks])
        assert_dag_equal(batched.to_numpy(), exp, f"segmented split {first_bits}")
        batched.free()
    finally:
        ctx.close()


def solid_keys(rng, depth, n_random):
    """A solid box (runs of all-ones leaves in Morton order) with a cavity and
    scattered voxels around it (distinct masks between the runs)."""
    side = 1 << depth
    lo, hi = side // 8, side // 8 + side // 4
    ax = np.arange(lo, hi, dtype=np.uint64)
    x, y, z = np.meshgrid(ax, ax, ax, indexing="ij")
    x, y, z = x.ravel(), y.ravel(), z.ravel()
    c = np.uint64(lo + side // 8)
    keep = (x - c) * (x - c) + (y - c) * (y - c) + (z - c) * (z - c) > np.uint64((side // 16) ** 2)
    x, y, z = x[keep], y[keep], z[keep]
    code = np.zeros(len(x), dtype=np.uint64)
    for b in range(depth):
        for a, v in enumerate((x, y, z)):
            code |= ((v >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b + 2 - a)
    extra = rng.integers(0, 1 << (3 * depth), size=n_random, dtype=np.uint64)
    keys = np.concatenate([code, extra, code[: len(code) // 5]])
    rng.shuffle(keys)
    return keys


@pytest.mark.parametrize("depth", [7, 9])
def test_solid_regions(gpu, oracle, depth):
    rng = np.random.default_rng(500 + depth)
    check_case(gpu, oracle, solid_keys(rng, depth, 3000), depth, f"solid d{depth}")
