"""Benchmark of the plain GPU DAG build (one MI355X).

Workload (default): triangle surfaces in a 2048^3 grid (depth 11): 20k seeded
triangles (v0 uniform in the grid, edges lognormal mu=4 sigma=1 voxels),
26-separating surface voxelization, 30% duplicates injected, shuffled uint64
Morton keys.  One step = svdag_create_dag + svdag_create_final_dag on keys
already resident in HBM (sort, leaf reduce, leaf and level dedup, per-level arrays
in ascending-hash order, flat uint32 layout), as DAGCompression::CreateDAG +
CreateFinalDAG do (src/DAGCompression.cu:1003-1107, 1550-1610).

Prints one JSON line: input voxels/s (keys including duplicates / step time),
the roofline of the dominant kernel (algorithmic bytes per launch / its average
launch time, measured with HIP events on the library's stream during the timed
steps), and the CPU oracle timed on the same keys on the host cores.
"""
import argparse
import ctypes
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "input voxels/s to bit-exact DAG, batched vs plain on 1 GPU; p99 chunk latency"

WORKLOADS = {
    "triangles2048": dict(desc="triangle surfaces 2048^3 (depth 11): 20k triangles, edge lognormal(4,1), "
                               "26-separating voxelization, 30% duplicates, shuffled u64 Morton keys",
                          depth=11, gen=("triangles", 0x5EED0001, 20000, 4.0, 1.0, 0.3)),
    "spheres256": dict(desc="sphere shells 256^3 (depth 8): 64 spheres r~U[8,40], 1-voxel shell, "
                            "~20% duplicates, shuffled u64 Morton keys",
                       depth=8, gen=("spheres", 0x5EED0000, 64, 8.0, 40.0, 0.25)),
}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_keys(svd, spec, threads):
    kind, seed, count, a, b, dup = spec["gen"]
    if kind == "triangles":
        return svd.gen_triangle_surfaces(seed, spec["depth"], count, a, b, dup, threads=threads)
    return svd.gen_sphere_shells(seed, spec["depth"], count, a, b, dup, threads=threads)


def algorithmic_bytes(kernel, st, key_bytes, launches, top_first=False):
    """Algorithmic HBM bytes one step moves through `kernel` (None if not modelled).

    N keys, L0 leaf positions, L unique leaves, M_l / Q_l nodes of level l before /
    after dedup, C_l children feeding level l, W words of the flat layout;
    `launches` = launches of the kernel per step (sort passes).  top_first: the
    key sort ran top digit first (its top pass writes 32-bit bucket-local keys,
    which the segmented passes and the leaf reduce read)."""
    n, l0, lu = st["n"], st["leaf_positions"], st["num_leaves"]
    pre, uniq, words = st["pre"], st["uniq"], st["words"]
    levels = len(pre)
    child = [pre[l + 1] if l + 1 < levels else l0 for l in range(levels)]
    sorted_bytes = 4 if top_first else key_bytes  # width of the keys the leaf reduce reads
    table = {
        # The first read of the keys (digit counts, range check).
        "sort_histogram": key_bytes * n,
        # LSD passes after the first count the keys they then scatter.
        "sort_count": launches * key_bytes * n,
        # Segmented counts read the 32-bit keys once per pass.
        "seg_count": launches * 4 * n if top_first else None,
        # Every LSD key-sort pass reads and writes each key once.
        "sort_scatter_keys": launches * 2 * key_bytes * n,
        # The top-digit pass reads each key and writes it as a 32-bit key.
        "sort_scatter_top": launches * (key_bytes + 4) * n,
        # Segmented passes after a top-digit pass read and write 32-bit keys.
        "seg_scatter": launches * 8 * n if top_first else None,
        # The unique-leaf sort moves (u64 mask, u32 slot) pairs; its first read
        # and its later passes' counts read the masks.
        "sort_scatter_pairs": launches * 2 * 12 * lu,
        "pair_histogram": 8 * lu,
        "pair_count": launches * 8 * lu,
        # grouped keys read once; (code, mask) per leaf into the chunk regions
        "leaf_reduce": sorted_bytes * n + 16 * l0,
        "leaf_compact": 32 * l0,
        # mask read and slot write per leaf position, table key per unique mask
        "leaf_insert": 12 * l0 + 8 * lu,
        # child code + slot + hash reads; parent code, slot, mask, 8 children, hash writes
        "level_build": sum(20 * child[l] + 53 * pre[l] for l in range(levels)),
        # rep + mask + children reads, rank + hash + mask + children writes per unique node
        "level_extract": sum(82 * uniq[l] for l in range(levels)),
        "final_write": sum(33 * uniq[l] for l in range(levels)) + 4 * sum(words) + 16 * lu,
    }
    return table.get(kernel)


def read_traffic(kernel, workload):
    """HBM bytes per launch from the committed rocprofv3 PMC summary, if there is one."""
    path = os.path.join(ROOT, "profiles", f"pmc_{kernel}_{workload}.md")
    if not os.path.exists(path):
        return None
    m = re.search(r"HBM bytes per launch \(scaled\):\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


def read_step_traffic(workload):
    """HBM bytes of one whole step (every kernel of create_dag + create_final_dag)
    from the committed rocprofv3 PMC summary, if there is one."""
    path = os.path.join(ROOT, "profiles", f"pmc_step_{workload}.md")
    if not os.path.exists(path):
        return None
    m = re.search(r"HBM bytes per step \(scaled\):\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


def device_words(L, _lib, ctx, dag):
    """Flat layout of a device DAG, left in device memory: (pointer, count)."""
    words, nw = ctypes.c_void_p(), ctypes.c_uint64()
    _lib.check(L.svdag_create_final_dag(ctx.handle, ctypes.byref(dag), ctypes.byref(words), ctypes.byref(nw)))
    return words, int(nw.value)


def fetch_words(L, _lib, ctx, words, nw):
    """Copies device words to the host and frees them."""
    out = np.empty(nw, dtype=np.uint32)
    _lib.hip_check(_lib.hip().hipMemcpy(out.ctypes.data, words, out.nbytes, _lib.HIP_D2H), "hipMemcpy")
    L.svdag_free_words(ctx.handle, words)
    return out


def final_words(L, _lib, ctx, dag):
    """Flat layout of a device DAG, copied to the host."""
    return fetch_words(L, _lib, ctx, *device_words(L, _lib, ctx, dag))


def bench_voxelize(svd, spec, args, ctx):
    """Device fragment producer on the triangle workload's own triangles:
    triangles (resident in HBM) -> unsorted Morton keys, then the DAG build of
    those keys.  Wall time per call, mean of `reps` after one warm-up."""
    kind, seed, count, mu, sigma, _ = spec["gen"]
    if kind != "triangles":
        return None
    depth = spec["depth"]
    tris = svd.gen_triangles(seed, depth, count, mu, sigma)
    d_tris = svd.DeviceBuffer.from_numpy(tris)
    L = svd._lib.lib()
    reps = 5

    def voxelize():
        p, n = ctypes.c_void_p(), ctypes.c_uint64()
        svd._lib.check(L.svdag_voxelize_triangles(ctx.handle, d_tris.ptr, count, depth, ctypes.byref(p),
                                                  ctypes.byref(n)))
        return p, int(n.value)

    p, n = voxelize()
    L.svdag_free_keys(ctx.handle, p)
    t0 = time.perf_counter()
    for _ in range(reps):
        p, n = voxelize()
        L.svdag_free_keys(ctx.handle, p)
    t_vox = (time.perf_counter() - t0) / reps
    # Triangles -> keys -> DAG -> flat layout, keys never leaving the device.
    t0 = time.perf_counter()
    for _ in range(reps):
        p, n = voxelize()
        dag = svd._lib.SvdagDag()
        svd._lib.check(L.svdag_create_dag(ctx.handle, p, n, depth, ctypes.byref(dag)))
        words, nw = ctypes.c_void_p(), ctypes.c_uint64()
        svd._lib.check(L.svdag_create_final_dag(ctx.handle, ctypes.byref(dag), ctypes.byref(words),
                                                ctypes.byref(nw)))
        L.svdag_free_words(ctx.handle, words)
        L.svdag_free_dag(ctx.handle, ctypes.byref(dag))
        L.svdag_free_keys(ctx.handle, p)
    t_e2e = (time.perf_counter() - t0) / reps
    if args.kernel_table:
        ctx.profile_enable(None)
        ctx.profile_reset()
        p, n = voxelize()
        L.svdag_free_keys(ctx.handle, p)
        table = ctx.profile_read()
        ctx.profile_enable([])
        for k, (ms, cnt) in sorted(table.items(), key=lambda kv: -kv[1][0]):
            log(f"  voxelize {k:20s} {ms:8.3f} ms  {cnt:4d} launches")
    d_tris.free()
    log(f"voxelize: {count} triangles -> {n} keys in {t_vox * 1e3:.2f} ms; with the build {t_e2e * 1e3:.2f} ms")
    return {"triangles": count, "keys": n, "voxelize_ms": round(t_vox * 1e3, 3),
            "keys_per_s": round(n / t_vox, 1), "triangles_to_words_ms": round(t_e2e * 1e3, 3)}


def bench_trace(svd, words, depth, args, ctx):
    """Primary rays through the step's flat layout (resident in HBM): one
    1920x1080 frame from a camera outside the volume looking at its centre,
    position colours.  Wall time per frame, mean of `reps` after one warm-up."""
    side = float(1 << depth)
    eye = np.array([-0.45, -0.35, 1.3]) * side
    fwd = np.full(3, 0.5 * side) - eye
    fwd /= np.linalg.norm(fwd)
    right = np.cross(fwd, [0.0, 0.0, 1.0])
    right /= np.linalg.norm(right)
    up = np.cross(right, fwd)
    width, height = 1920, 1080
    p = svd.trace_params(eye, right, up, fwd, 60.0, width, height, (0.0, 0.0, 0.0), (side, side, side), depth,
                         debug_level=0xFFFFFFFF)
    tp = svd._lib.TraceParams.from_dict(p)
    L = svd._lib.lib()
    dw = svd.DeviceBuffer.from_numpy(words)
    dpaths = svd.DeviceBuffer(12 * width * height)
    dcolors = svd.DeviceBuffer(4 * width * height)
    reps = 10

    def frame():
        svd._lib.check(L.svdag_trace_paths(ctx.handle, dw.ptr, words.size, depth, ctypes.byref(tp), dpaths.ptr,
                                           dcolors.ptr))

    try:
        frame()
        t0 = time.perf_counter()
        for _ in range(reps):
            frame()
        dt = (time.perf_counter() - t0) / reps
        hits = int(dpaths.to_numpy(np.uint32, 3 * width * height).reshape(-1, 3).any(axis=1).sum())
        ctx.profile_enable(["trace_paths"])
        ctx.profile_reset()
        frame()
        kernel_ms = ctx.profile_read()["trace_paths"][0]
        ctx.profile_enable([])
    finally:
        for b in (dw, dpaths, dcolors):
            b.free()
    log(f"trace: {width}x{height} in {dt * 1e3:.2f} ms ({hits} rays hit), kernel {kernel_ms:.3f} ms")
    return {"width": width, "height": height, "frame_ms": round(dt * 1e3, 3), "kernel_ms": round(kernel_ms, 3),
            "rays_per_s": round(width * height / dt, 1), "hits": hits}


# The reference's recorded sort throughputs (100M random keys on a GTX
# 1080-class GPU): 64-bit keys on 33 bits (Python/variable64.csv:33) and 64
# bits (constant64.csv:15), 32-bit keys on 27 bits (variable32.csv:27): the
# only numbers it publishes for a stage of this path.
REFERENCE_SORT = {33: 1.751e9, 64: 0.925e9, "u32_27": 4.167e9}


def bench_sort(svd, args, ctx):
    """svdag_sort_keys (the reference's SortKeys) on 100M random uint64 keys,
    device-resident, at 33 bits (depth 11 Morton keys) and 64 bits.  Wall time
    per call, best of `reps` after one warm-up."""
    n, reps = 100_000_000, 5
    rng = np.random.default_rng(0x5EED0005)
    out = {}
    d_out = svd.DeviceBuffer(8 * n)
    L = svd._lib.lib()
    try:
        for name in (33, 64, "u32_27"):
            bits = 27 if name == "u32_27" else name
            hi = 1 << bits if bits < 64 else None
            keys = (rng.integers(0, hi, size=n, dtype=np.uint64) if hi
                    else rng.integers(0, 2 ** 63, size=n, dtype=np.uint64) << np.uint64(1))
            fn = L.svdag_sort_keys
            if name == "u32_27":
                keys = keys.astype(np.uint32)
                fn = L.svdag_sort_keys_u32
            d_keys = svd.DeviceBuffer.from_numpy(keys)
            try:
                best = None
                for r in range(reps + 1):
                    t0 = time.perf_counter()
                    svd._lib.check(fn(ctx.handle, d_keys.ptr, n, bits, d_out.ptr))
                    dt = time.perf_counter() - t0
                    if r:
                        best = dt if best is None else min(best, dt)
                ok = np.array_equal(d_out.to_numpy(keys.dtype, 1 << 20), np.sort(keys)[: 1 << 20])
            finally:
                d_keys.free()
            rate = n / best
            out[str(name)] = {"ms": round(best * 1e3, 3), "keys_per_s": round(rate, 1),
                              "reference_keys_per_s": REFERENCE_SORT[name],
                              "vs_reference": round(rate / REFERENCE_SORT[name], 1), "prefix_checked": bool(ok)}
            log(f"sort_keys {name} bits: {n} keys in {best * 1e3:.2f} ms = {rate / 1e9:.1f} G keys/s "
                f"({rate / REFERENCE_SORT[name]:.0f}x the reference's recorded rate)")
    finally:
        d_out.free()
    return {"n": n, "by_bits": out}


def bench_chunks(svd, _lib, args):
    """Plain (CreateDAG per chunk + MergeDAGs + CreateFinalDAG, the reference's
    flow), one plain build of the union of the chunks, batched (one
    create_dag_batched) and served (continuous admission) on the same 512
    uneven chunks; all must give the oracle's DAG of the union."""
    from dag_compression_fast_amd import serve
    from dag_compression_fast_amd import workloads as W
    L, H = _lib.lib(), _lib.hip()
    spec = W.CHUNKS
    chunks = W.chunk_workload(args.threads)
    sizes = np.array([len(c) for c in chunks], dtype=np.uint64)
    offsets = W.chunk_offsets(chunks)
    total = int(offsets[-1])
    log(f"chunks: {len(chunks)} chunks, {total} keys")
    d_keys = svd.DeviceBuffer.from_numpy(np.concatenate(chunks))
    ctx = svd.Context()
    depth = spec["depth"]
    global_depth = depth + spec["levels_above"]

    def plain():
        dags = (_lib.SvdagDag * len(chunks))()
        for i in range(len(chunks)):
            ptr = ctypes.c_void_p(d_keys.address + 8 * int(offsets[i]))
            _lib.check(L.svdag_create_dag(ctx.handle, ptr, int(sizes[i]), depth, ctypes.byref(dags[i])))
        out = _lib.SvdagDag()
        _lib.check(L.svdag_merge_dags(ctx.handle, dags, len(chunks), ctypes.byref(out)))
        for i in range(len(chunks)):
            L.svdag_free_dag(ctx.handle, ctypes.byref(dags[i]))
        return out

    def batched():
        out = _lib.SvdagDag()
        _lib.check(L.svdag_create_dag_batched(ctx.handle, d_keys.ptr,
                                              offsets.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                              len(chunks), depth, ctypes.byref(out)))
        return out

    def timed(fn, reps):
        """Best of `reps` runs after one warm-up: build + flat layout from device-resident
        keys to device-resident words (the copy to the host for the check is outside)."""
        best, words = None, None
        for r in range(reps + 1):
            _lib.hip_check(H.hipDeviceSynchronize(), "hipDeviceSynchronize")
            t = time.perf_counter()
            dag = fn()
            d_words, nw = device_words(L, _lib, ctx, dag)
            dt = time.perf_counter() - t
            words = fetch_words(L, _lib, ctx, d_words, nw)
            if r > 0:
                best = dt if best is None else min(best, dt)
            L.svdag_free_dag(ctx.handle, ctypes.byref(dag))
        return best, words

    t_plain, w_plain = timed(plain, args.chunk_reps)
    log(f"chunks plain: {t_plain * 1e3:.1f} ms")
    # One plain build of the union (global codes at depth 12) from device-resident keys.
    union = W.chunk_union(chunks, depth)
    d_union = svd.DeviceBuffer.from_numpy(union)

    def union_build():
        out = _lib.SvdagDag()
        _lib.check(L.svdag_create_dag(ctx.handle, d_union.ptr, total, global_depth, ctypes.byref(out)))
        return out

    t_union, w_union = timed(union_build, args.chunk_reps)
    d_union.free()
    log(f"chunks union, one plain build: {t_union * 1e3:.2f} ms")
    tuned = None
    if not args.no_autotune:
        # Table load factor and the segmented sort's digit split for the
        # batched build, picked by timing it (every candidate must give the same DAG).
        from dag_compression_fast_amd import autotune
        state = {}

        def tune_build(c):
            _lib.hip_check(H.hipDeviceSynchronize(), "hipDeviceSynchronize")
            dag = batched()
            if "words" in state:
                L.svdag_free_words(ctx.handle, state.pop("words")[0])
            state["words"] = device_words(L, _lib, ctx, dag)
            L.svdag_free_dag(ctx.handle, ctypes.byref(dag))

        def tune_fetch():
            return fetch_words(L, _lib, ctx, *state.pop("words"))

        t0 = time.perf_counter()
        res = autotune.autotune(ctx, tune_build, tune_fetch, reps=3, space=autotune.space_for(depth),
                                chunk_depth=depth)
        if "words" in state:
            L.svdag_free_words(ctx.handle, state.pop("words")[0])
        times = [ms for _, ms in res["table"]]
        tuned = {"best": res["best"], "best_ms": round(res["best_ms"], 3), "candidates": len(res["table"]),
                 "default_ms": round(res["default_ms"], 3), "worst_ms": round(max(times), 3),
                 "tune_s": round(time.perf_counter() - t0, 2)}
        log(f"autotune: {res['best']} {res['best_ms']:.2f} ms (default {tuned['default_ms']:.2f} ms, worst "
            f"{tuned['worst_ms']:.2f} ms) over {len(res['table'])} settings in {tuned['tune_s']:.1f} s")
    t_batch, w_batch = timed(batched, args.chunk_reps)
    log(f"chunks batched: {t_batch * 1e3:.2f} ms")
    # The same build captured once as a HIP graph and replayed (offsets in
    # device memory, no host planning per replay).
    d_off = svd.DeviceBuffer.from_numpy(offsets)
    graph = svd.BatchedGraph(d_keys, d_off, len(chunks), depth, total, ctx=ctx)

    def replay():
        graph.launch()
        d = graph.result()
        raw, d.raw = d.raw, None  # the arrays live in the graph; nothing to free
        return raw

    try:
        t_graph, w_graph = timed(replay, args.chunk_reps)
    finally:
        graph.close()
        d_off.free()
    log(f"chunks batched, graph replay: {t_graph * 1e3:.2f} ms")
    if args.kernel_table:
        ctx.profile_enable(None)
        ctx.profile_reset()
        L.svdag_free_dag(ctx.handle, ctypes.byref(batched()))
        table = ctx.profile_read()
        ctx.profile_enable([])
        for k, (ms, cnt) in sorted(table.items(), key=lambda kv: -kv[1][0])[:12]:
            log(f"  batched {k:20s} {ms:8.3f} ms  {cnt:4d} launches")
    same = np.array_equal(w_plain, w_batch) and np.array_equal(w_union, w_batch) and np.array_equal(w_graph, w_batch)

    served = bench_served(svd, _lib, args, d_keys, offsets, w_batch)
    same = same and served["saturated"]["bit_exact_vs_batched"] and all(
        r["same_dag_as_other_rates"] and r["bit_exact_vs_merge"] for r in served.get("sweep", {}).get("rates", []))
    d_keys.free()
    ctx.close()

    check = "plain, union, batched, graph and served final DAGs identical" if same else "MISMATCH"
    cpu = None
    if not args.no_cpu_baseline and same:
        # The oracle on the union: the CPU baseline of this workload and the check.
        log(f"chunks: CPU oracle on the union ({total} keys), {args.threads} threads")
        t0 = time.perf_counter()
        ora = oracle_build(union, global_depth, args.threads)
        t_cpu = time.perf_counter() - t0
        if np.array_equal(ora, w_batch):
            check = ("bit-exact vs oracle (union of the chunks at depth 12; plain, union, batched, graph and served "
                     "identical)")
        else:
            check = "MISMATCH vs oracle"
        cpu = {"value": round(total / t_cpu, 1), "unit": "voxels/s", "cores": args.threads, "kind": "port",
               "sample": "full workload, one run: create_dag of the union + create_final_dag"}
    return {"workload": W.CHUNKS["desc"], "n_keys": total, "plain_voxels_per_s": round(total / t_plain, 1),
            "plain_ms": round(t_plain * 1e3, 2), "union_voxels_per_s": round(total / t_union, 1),
            "union_ms": round(t_union * 1e3, 3), "batched_voxels_per_s": round(total / t_batch, 1),
            "batched_ms": round(t_batch * 1e3, 3), "batched_vs_plain": round(t_plain / t_batch, 2),
            "graph_ms": round(t_graph * 1e3, 3), "graph_voxels_per_s": round(total / t_graph, 1),
            "batched_vs_union": round(t_union / t_batch, 3),
            "served": served, "served_vs_batched": round(served["saturated"]["voxels_per_s"] * t_batch / total, 3),
            "autotune": tuned, "cpu_baseline": cpu, "check": check}


def bench_served(svd, _lib, args, d_keys, offsets, w_batch):
    """Served mode (continuous batching into a resident global DAG).

    Saturated: the 512 chunks already in device memory all arrive at once;
    rounds are cut by the HBM budget; the final DAG must equal the batched
    build's.  Sweep: 2000 requests drawn as the 512-chunk workload at distinct
    slots of a 4096-slot grid, Poisson arrivals at each rate (2000 requests, or
    30 s at the lowest rates); every rate must end with the same global DAG,
    which must equal the reference's flow over the same chunks (8 batched
    builds of 512 slots merged by MergeDAGs)."""
    from dag_compression_fast_amd import serve
    from dag_compression_fast_amd import workloads as W
    L, H = _lib.lib(), _lib.hip()
    out = {}
    total = int(offsets[-1])
    views = [svd.DeviceView(d_keys, 8 * int(offsets[i]), int(offsets[i + 1] - offsets[i]))
             for i in range(len(offsets) - 1)]
    # One context for every server: a running server's scratch and result
    # blocks are warm.  The first saturated run warms them; the second is measured.
    sctx = svd.Context()
    for run in range(2):
        server = serve.ChunkServer(W.CHUNKS["depth"], W.CHUNKS["levels_above"], max_nodes=max(1 << 20, total // 16),
                                   ctx=sctx)
        try:
            reqs = [serve.Request(i, i, views[i], 0.0) for i in range(len(views))]
            results = server.run(reqs)
            span = max(r.done for r in results) - min(r.arrival for r in results)
            dag = server.dag()
            same = np.array_equal(final_words(L, _lib, sctx, dag.raw), w_batch)
            L.svdag_free_dag(sctx.handle, ctypes.byref(dag.raw))
            rounds = server.rounds
        finally:
            server.close()
    stats = serve.latency_stats(results)
    out["saturated"] = {"requests": len(reqs), "rounds": rounds, "voxels_per_s": round(total / span, 1),
                        "ms": round(span * 1e3, 3), **{k: round(v, 3) for k, v in stats.items()},
                        "bit_exact_vs_batched": bool(same)}
    log(f"served saturated: {len(reqs)} requests in {rounds} rounds, {span * 1e3:.2f} ms = "
        f"{total / span / 1e9:.1f} G voxels/s, identical to batched: {same}")
    if not args.serve_rates:
        sctx.close()
        return out

    spec = W.SERVED
    t0 = time.perf_counter()
    chunks = W.chunk_workload(args.threads, seed=spec["seed"], num=spec["requests"], depth=spec["depth"])
    sizes = np.array([len(c) for c in chunks], dtype=np.uint64)
    c_off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    all_keys = svd.DeviceBuffer.from_numpy(np.concatenate(chunks))
    del chunks
    nslots = 8 ** spec["slot_levels"]
    slot_of = np.random.default_rng(spec["seed"]).permutation(nslots)[: spec["requests"]]
    views = [svd.DeviceView(all_keys, 8 * int(c_off[i]), int(sizes[i])) for i in range(spec["requests"])]
    log(f"served sweep: {spec['requests']} requests, {int(c_off[-1])} keys, {nslots} slots "
        f"(generated and uploaded in {time.perf_counter() - t0:.1f} s)")
    sweep, words_seen = [], {}
    try:
        for rate in args.serve_rates:
            n_req = min(spec["requests"], max(1, int(np.ceil(spec["min_seconds"] * rate))))
            arrivals = serve.poisson_arrivals(rate, n_req, seed=int(rate))
            reqs = [serve.Request(i, int(slot_of[i]), views[i], float(arrivals[i])) for i in range(n_req)]
            keys = int(sizes[:n_req].sum())
            server = serve.ChunkServer(spec["depth"], spec["slot_levels"], max_nodes=max(1 << 20, keys // 16),
                                       ctx=sctx)
            try:
                t = time.perf_counter()
                results = server.run(reqs)
                span = max(r.done for r in results) - min(r.arrival for r in results)
                dag = server.dag()
                words = final_words(L, _lib, sctx, dag.raw)
                L.svdag_free_dag(sctx.handle, ctypes.byref(dag.raw))
                rounds = server.rounds
            finally:
                server.close()
            stats = serve.latency_stats(results)
            key = n_req
            if key not in words_seen:
                words_seen[key] = words
            same = np.array_equal(words, words_seen[key])
            sweep.append({"rate_per_s": rate, "requests": n_req, "seconds": round(span, 2), "rounds": rounds,
                          "voxels_per_s": round(keys / span, 1), **{k: round(v, 3) for k, v in stats.items()},
                          "same_dag_as_other_rates": bool(same)})
            log(f"served {rate:.0f}/s: {n_req} requests, {rounds} rounds, {keys / span / 1e9:.2f} G voxels/s, "
                f"p50 {stats['p50_ms']:.2f} ms p99 {stats['p99_ms']:.2f} ms ({time.perf_counter() - t:.1f} s)")
        # The reference's flow over the same chunks: the 8 octants of the
        # slot grid (512 slots each, batched) merged by MergeDAGs.
        ref_words = {}
        per = nslots // 8
        for n_req in sorted(words_seen):
            by_slot = {int(slot_of[i]): views[i] for i in range(n_req)}
            dags = (_lib.SvdagDag * 8)()
            ctx = svd.Context()
            try:
                for o in range(8):
                    bufs = [by_slot.get(o * per + j) for j in range(per)]
                    d = svd.create_dag_batched_gather(bufs, spec["depth"], ctx=ctx)
                    dags[o] = d.raw
                    d.raw = _lib.SvdagDag()  # ownership moves to the array
                merged = _lib.SvdagDag()
                _lib.check(L.svdag_merge_dags(ctx.handle, dags, 8, ctypes.byref(merged)))
                for o in range(8):
                    L.svdag_free_dag(ctx.handle, ctypes.byref(dags[o]))
                ref_words[n_req] = final_words(L, _lib, ctx, merged)
                L.svdag_free_dag(ctx.handle, ctypes.byref(merged))
            finally:
                ctx.close()
        for row in sweep:
            row["bit_exact_vs_merge"] = bool(np.array_equal(words_seen[row["requests"]], ref_words[row["requests"]]))
    finally:
        all_keys.free()
        sctx.close()
    out["sweep"] = {"workload": spec["desc"], "rates": sweep}
    return out


def bench_noise(svd, _lib, args):
    """The low-redundancy value-noise grid (1024^3, depth 10, 8 batches of
    512^3 built on their own and merged, the reference's many-chunk flow) and
    the 512^3 cube with 32-bit keys; keys device-resident, results ending as
    device-resident words.  Checked against the oracle in the CPU-baseline leg."""
    from dag_compression_fast_amd import workloads as W
    L, H = _lib.lib(), _lib.hip()
    spec = W.NOISE
    t0 = time.perf_counter()
    batches = W.noise_batches(args.threads)
    sizes = [len(b) for b in batches]
    total = sum(sizes)
    log(f"noise1024: {total} keys in {len(batches)} batches (generated in {time.perf_counter() - t0:.1f} s)")
    sub = spec["batch_depth"]
    d_keys = svd.DeviceBuffer.from_numpy(np.concatenate(batches))
    offsets = np.zeros(len(batches) + 1, dtype=np.uint64)
    offsets[1:] = np.cumsum(sizes)
    ctx = svd.Context()

    def per_batch():
        dags = (_lib.SvdagDag * len(batches))()
        for i in range(len(batches)):
            ptr = ctypes.c_void_p(d_keys.address + 8 * int(offsets[i]))
            _lib.check(L.svdag_create_dag(ctx.handle, ptr, sizes[i], sub, ctypes.byref(dags[i])))
        out = _lib.SvdagDag()
        _lib.check(L.svdag_merge_dags(ctx.handle, dags, len(batches), ctypes.byref(out)))
        for i in range(len(batches)):
            L.svdag_free_dag(ctx.handle, ctypes.byref(dags[i]))
        return out

    def batched():
        out = _lib.SvdagDag()
        _lib.check(L.svdag_create_dag_batched(ctx.handle, d_keys.ptr,
                                              offsets.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                              len(batches), sub, ctypes.byref(out)))
        return out

    def timed(fn, reps):
        best, words, counts = None, None, None
        for r in range(reps + 1):
            _lib.hip_check(H.hipDeviceSynchronize(), "hipDeviceSynchronize")
            t = time.perf_counter()
            dag = fn()
            d_words, nw = device_words(L, _lib, ctx, dag)
            dt = time.perf_counter() - t
            words = fetch_words(L, _lib, ctx, d_words, nw)
            counts = [int(dag.levels[l].num_nodes) for l in range(dag.num_levels)] + [int(dag.num_leaves)]
            if r > 0:
                best = dt if best is None else min(best, dt)
            L.svdag_free_dag(ctx.handle, ctypes.byref(dag))
        return best, words, counts

    t_plain, w_plain, counts = timed(per_batch, args.chunk_reps)
    t_batch, w_batch, _ = timed(batched, args.chunk_reps)
    d_keys.free()
    log(f"noise1024: 8 builds + merge {t_plain * 1e3:.2f} ms, batched {t_batch * 1e3:.2f} ms; "
        f"nodes per level {counts}")

    # 512^3 with 32-bit keys, and the same keys as 64-bit keys.
    keys32 = W.noise_u32(args.threads)
    n32 = len(keys32)
    d32 = svd.DeviceBuffer.from_numpy(keys32)
    d64 = svd.DeviceBuffer.from_numpy(keys32.astype(np.uint64))

    def u32():
        out = _lib.SvdagDag()
        _lib.check(L.svdag_create_dag_u32(ctx.handle, d32.ptr, n32, W.NOISE_U32["depth"], ctypes.byref(out)))
        return out

    def u64():
        out = _lib.SvdagDag()
        _lib.check(L.svdag_create_dag(ctx.handle, d64.ptr, n32, W.NOISE_U32["depth"], ctypes.byref(out)))
        return out

    t32, w32, _ = timed(u32, args.chunk_reps)
    t64, w64, _ = timed(u64, args.chunk_reps)
    d32.free()
    d64.free()
    log(f"noise512: u32 keys {t32 * 1e3:.2f} ms, u64 keys {t64 * 1e3:.2f} ms ({n32} keys)")
    same = np.array_equal(w_plain, w_batch) and np.array_equal(w32, w64)
    check = "per-batch + merge and batched identical; u32 and u64 identical" if same else "MISMATCH"
    cpu = None
    if not args.no_cpu_baseline and same:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_ctypes
        log(f"noise1024: CPU oracle, 8 builds + merge, {args.threads} threads")
        t0 = time.perf_counter()
        ora = oracle_ctypes.create_final_dag(oracle_ctypes.merge_dags(
            [oracle_ctypes.create_dag(b, sub, threads=args.threads) for b in batches], threads=args.threads))
        t_cpu = time.perf_counter() - t0
        ora32 = oracle_build(keys32, W.NOISE_U32["depth"], args.threads)
        ok = np.array_equal(ora, w_plain) and np.array_equal(ora32, w32)
        check = ("bit-exact vs oracle (1024^3: 8 batch builds merged; 512^3 u32)" if ok else "MISMATCH vs oracle")
        cpu = {"value": round(total / t_cpu, 1), "unit": "voxels/s", "cores": args.threads, "kind": "port",
               "sample": "full 1024^3 workload, one run: 8 create_dag + merge_dags + create_final_dag"}
    ctx.close()
    return {"workload": spec["desc"], "n_keys": total, "nodes_per_level": counts,
            "plain_ms": round(t_plain * 1e3, 3), "plain_voxels_per_s": round(total / t_plain, 1),
            "batched_ms": round(t_batch * 1e3, 3), "batched_voxels_per_s": round(total / t_batch, 1),
            "u32": {"workload": W.NOISE_U32["desc"], "n_keys": n32, "ms": round(t32 * 1e3, 3),
                    "voxels_per_s": round(n32 / t32, 1), "u64_ms": round(t64 * 1e3, 3)},
            "cpu_baseline": cpu, "check": check}


def oracle_build(keys, depth, threads):
    """Oracle CreateDAG + CreateFinalDAG of host keys: the flat words."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_ctypes
    return oracle_ctypes.create_final_dag(oracle_ctypes.create_dag(keys, depth, threads=threads))


def profile_batched(svd, _lib, args):
    """N batched builds (+ flat layout) of the chunk workload and nothing else,
    so that a profiler's kernel statistics cover exactly that path."""
    from dag_compression_fast_amd import workloads as W
    L = _lib.lib()
    chunks = W.chunk_workload(args.threads)
    offsets = W.chunk_offsets(chunks)
    d_keys = svd.DeviceBuffer.from_numpy(np.concatenate(chunks))
    ctx = svd.Context()
    times = []
    try:
        for _ in range(args.profile_batched):
            t0 = time.perf_counter()
            out = _lib.SvdagDag()
            _lib.check(L.svdag_create_dag_batched(ctx.handle, d_keys.ptr,
                                                  offsets.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                                  len(chunks), W.CHUNKS["depth"], ctypes.byref(out)))
            words, nw = device_words(L, _lib, ctx, out)
            times.append(time.perf_counter() - t0)
            L.svdag_free_words(ctx.handle, words)
            L.svdag_free_dag(ctx.handle, ctypes.byref(out))
    finally:
        d_keys.free()
        ctx.close()
    print(json.dumps({"batched_builds": len(times), "n_keys": int(offsets[-1]),
                      "best_ms": round(min(times) * 1e3, 3)}), flush=True)
    return 0


def profile_served(svd, _lib, args):
    """N saturated served runs of the 512-chunk workload (every chunk arrives at
    once; rounds cut by the HBM budget), for a profiler; the first run warms
    the shared context."""
    from dag_compression_fast_amd import serve
    from dag_compression_fast_amd import workloads as W
    chunks = W.chunk_workload(args.threads)
    offsets = W.chunk_offsets(chunks)
    d_keys = svd.DeviceBuffer.from_numpy(np.concatenate(chunks))
    views = [svd.DeviceView(d_keys, 8 * int(offsets[i]), len(chunks[i])) for i in range(len(chunks))]
    ctx = svd.Context()
    times, rounds = [], []
    try:
        for _ in range(args.profile_served):
            server = serve.ChunkServer(W.CHUNKS["depth"], W.CHUNKS["levels_above"],
                                       max_nodes=max(1 << 20, int(offsets[-1]) // 16), ctx=ctx)
            try:
                results = server.run([serve.Request(i, i, views[i], 0.0) for i in range(len(views))])
                times.append(max(r.done for r in results))
                rounds.append(server.rounds)
            finally:
                server.close()
    finally:
        d_keys.free()
        ctx.close()
    print(json.dumps({"served_runs": len(times), "rounds": rounds, "best_ms": round(min(times) * 1e3, 3)}), flush=True)
    return 0


def profile_noise(svd, _lib, args):
    """N rounds of the value-noise flow (8 batch builds + merge + flat layout)
    and nothing else, for a profiler: the dedup kernels at worst-case load."""
    from dag_compression_fast_amd import workloads as W
    L = _lib.lib()
    batches = W.noise_batches(args.threads)
    sizes = [len(b) for b in batches]
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    d_keys = svd.DeviceBuffer.from_numpy(np.concatenate(batches))
    ctx = svd.Context()
    times = []
    try:
        for _ in range(args.profile_noise):
            t0 = time.perf_counter()
            dags = (_lib.SvdagDag * len(batches))()
            for i in range(len(batches)):
                ptr = ctypes.c_void_p(d_keys.address + 8 * int(offsets[i]))
                _lib.check(L.svdag_create_dag(ctx.handle, ptr, sizes[i], W.NOISE["batch_depth"],
                                              ctypes.byref(dags[i])))
            out = _lib.SvdagDag()
            _lib.check(L.svdag_merge_dags(ctx.handle, dags, len(batches), ctypes.byref(out)))
            words, nw = device_words(L, _lib, ctx, out)
            times.append(time.perf_counter() - t0)
            L.svdag_free_words(ctx.handle, words)
            L.svdag_free_dag(ctx.handle, ctypes.byref(out))
            for i in range(len(batches)):
                L.svdag_free_dag(ctx.handle, ctypes.byref(dags[i]))
    finally:
        d_keys.free()
        ctx.close()
    print(json.dumps({"noise_rounds": len(times), "n_keys": int(offsets[-1]),
                      "best_ms": round(min(times) * 1e3, 3)}), flush=True)
    return 0


def cpu_baseline(keys, depth, threads, reps):
    """The CPU oracle (a restatement of the reference algorithm) on the same keys."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_ctypes
    oracle_ctypes.create_dag(keys, depth, threads=threads)  # warm-up
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        d = oracle_ctypes.create_dag(keys, depth, threads=threads)
        oracle_ctypes.create_final_dag(d)
        times.append(time.perf_counter() - t0)
    return float(np.median(times)), d


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="triangles2048", choices=sorted(WORKLOADS))
    ap.add_argument("--threads", type=int, default=min(16, os.cpu_count() or 1),
                    help="host threads for input generation and the CPU baseline")
    ap.add_argument("--cpu-reps", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--kernel-table", action="store_true", help="print per-kernel times to stderr")
    ap.add_argument("--radix-bits", type=int, default=None, help="widest key-sort digit (library default if unset)")
    ap.add_argument("--sort-tile", type=int, default=None, help="keys per key-sort tile, 4096 or 8192")
    ap.add_argument("--top-digit-first", type=int, default=None, choices=(0, 1, 2),
                    help="key sort route of plain 64-bit builds: 0 LSD, 1 auto (library default), 2 top digit first")
    ap.add_argument("--no-chunks", action="store_true", help="skip the 512-chunk batched/served comparison")
    ap.add_argument("--no-noise", action="store_true", help="skip the value-noise workloads")
    ap.add_argument("--no-voxelize", action="store_true", help="skip the device voxeliser timing")
    ap.add_argument("--no-trace", action="store_true", help="skip the ray tracing timing")
    ap.add_argument("--no-autotune", action="store_true", help="batched mode with the default options")
    ap.add_argument("--no-sort", action="store_true", help="skip the sort_keys timing")
    ap.add_argument("--profile-batched", type=int, default=0, metavar="N",
                    help="only run N batched builds of the chunk workload (for a profiler), then exit")
    ap.add_argument("--profile-served", type=int, default=0, metavar="N",
                    help="only run N saturated served runs of the chunk workload (for a profiler), then exit")
    ap.add_argument("--profile-noise", type=int, default=0, metavar="N",
                    help="only run N rounds of the value-noise flow (for a profiler), then exit")
    ap.add_argument("--chunk-reps", type=int, default=2)
    ap.add_argument("--serve-rates", type=float, nargs="*", default=[50.0, 100.0, 200.0, 500.0, 1000.0, 2000.0],
                    help="Poisson arrival rates (requests/s) of the served sweep (none: saturated run only)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (the build runs on a single "
              "MI355X; sharding across GPUs is a design note only)", file=sys.stderr)
        sys.exit(2)

    import dag_compression_fast_amd as svd
    from dag_compression_fast_amd import _lib

    if _lib.gpu_count() < 1:
        sys.exit("bench.py needs a visible GPU")
    if args.profile_batched:
        return profile_batched(svd, _lib, args)
    if args.profile_noise:
        return profile_noise(svd, _lib, args)
    if args.profile_served:
        return profile_served(svd, _lib, args)
    spec = WORKLOADS[args.workload]
    depth = spec["depth"]
    log(f"generating {args.workload} on {args.threads} host threads")
    keys = make_keys(svd, spec, args.threads)
    n = len(keys)
    log(f"{n} keys, depth {depth}")

    ctx = svd.Context()
    if args.radix_bits is not None:
        ctx.set_option(_lib.OPT_RADIX_BITS, args.radix_bits)
    if args.sort_tile is not None:
        ctx.set_option(_lib.OPT_SORT_TILE, args.sort_tile)
    if args.top_digit_first is not None:
        ctx.set_option(_lib.OPT_TOP_DIGIT_FIRST, args.top_digit_first)
    L = _lib.lib()
    d_keys = svd.DeviceBuffer.from_numpy(keys)
    H = _lib.hip()

    def step():
        dag = _lib.SvdagDag()
        _lib.check(L.svdag_create_dag(ctx.handle, d_keys.ptr, n, depth, ctypes.byref(dag)))
        words, nw = ctypes.c_void_p(), ctypes.c_uint64()
        _lib.check(L.svdag_create_final_dag(ctx.handle, ctypes.byref(dag), ctypes.byref(words), ctypes.byref(nw)))
        return dag, words, int(nw.value)

    def release(dag, words):
        L.svdag_free_words(ctx.handle, words)
        L.svdag_free_dag(ctx.handle, ctypes.byref(dag))

    # Warm-up, with every kernel timed once to find the dominant one.
    ctx.profile_enable(None)
    for i in range(max(1, args.warmup)):
        if i == max(1, args.warmup) - 1:
            ctx.profile_reset()
        dag, words, nw = step()
        if i < max(1, args.warmup) - 1:
            release(dag, words)
    table = ctx.profile_read()
    stats = {
        "n": n, "leaf_positions": int(dag.leaf_positions), "num_leaves": int(dag.num_leaves),
        "pre": [int(dag.pre_dedup[l]) for l in range(dag.num_levels)],
        "uniq": [int(dag.levels[l].num_nodes) for l in range(dag.num_levels)],
        "words": [int(dag.level_words[l]) for l in range(dag.num_levels)],
    }
    result_words = np.empty(nw, dtype=np.uint32)
    _lib.hip_check(H.hipMemcpy(result_words.ctypes.data, words, 4 * nw, _lib.HIP_D2H), "hipMemcpy")
    level_counts = stats["uniq"]
    release(dag, words)
    dominant = max(table, key=lambda k: table[k][0])
    if args.kernel_table:
        tot = sum(v[0] for v in table.values())
        for k, (ms, cnt) in sorted(table.items(), key=lambda kv: -kv[1][0]):
            log(f"  {k:20s} {ms:8.3f} ms  {cnt:4d} launches  {100 * ms / tot:5.1f}%")
    log(f"dominant kernel: {dominant}")

    # Timed steps.  The dominant kernel carries timing events in the last timed
    # step only: an event pair around each of its launches costs about 9 us of
    # idle GPU time (44 us per c1 step with 5 launches), which timing every
    # step would add to the step time it measures.
    ctx.profile_enable([])
    ctx.profile_reset()
    _lib.hip_check(H.hipDeviceSynchronize(), "hipDeviceSynchronize")
    t0 = time.perf_counter()
    for i in range(args.steps):
        if i == args.steps - 1:
            ctx.profile_enable([dominant])
        dag, words, _nw = step()
        release(dag, words)
    _lib.hip_check(H.hipDeviceSynchronize(), "hipDeviceSynchronize")
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    kms, launches = ctx.profile_read()[dominant]
    ctx.profile_enable([])

    launches_per_step = launches  # the one profiled step
    top_first = "sort_scatter_top" in table
    alg = algorithmic_bytes(dominant, stats, 8, launches_per_step, top_first)
    roofline = None
    if alg is not None:
        per_launch = alg / launches_per_step
        avg_s = kms / launches / 1e3
        achieved = per_launch / avg_s / 1e9
        traffic = read_traffic(dominant, args.workload)
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": dominant,
                    "bytes_per_launch": int(per_launch), "avg_launch_us": round(avg_s * 1e6, 2)}
    # The whole step against the HBM roofline: algorithmic bytes of every
    # modelled kernel (one warm-up step's launches), and the counter-measured
    # bytes of a step where a committed PMC summary holds them.
    step_alg = 0.0
    for k, (_ms, cnt) in table.items():
        b = algorithmic_bytes(k, stats, 8, cnt, top_first)
        if b is not None:
            step_alg += b
    step_traffic = read_step_traffic(args.workload)
    step_s = ms_per_step / 1e3
    whole_step = {"algorithmic_bytes": int(step_alg), "frac": round(step_alg / step_s / 1e9 / HBM_PEAK_GBS, 4),
                  "traffic": step_traffic,
                  "traffic_frac": round(step_traffic / step_s / 1e9 / HBM_PEAK_GBS, 4) if step_traffic else None}
    if roofline is not None:
        roofline["whole_step"] = whole_step

    cpu = None
    check = "not checked"
    if not args.no_cpu_baseline:
        log(f"CPU oracle on the same {n} keys, {args.threads} threads")
        cpu_s, ref = cpu_baseline(keys, depth, args.threads, args.cpu_reps)
        cpu = {"value": round(n / cpu_s, 1), "unit": "voxels/s", "cores": args.threads, "kind": "port",
               "sample": f"full workload ({n} keys), median of {args.cpu_reps} after 1 warm-up, "
                         "create_dag + create_final_dag"}
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_ctypes
        ref_counts = [len(lv["hashes"]) for lv in ref["levels"]]
        same = ref_counts == level_counts and np.array_equal(oracle_ctypes.create_final_dag(ref), result_words)
        check = "bit-exact vs oracle (level counts + flat layout)" if same else "MISMATCH vs oracle"
        if not same:
            log("GPU result differs from the oracle")

    # The reference's contract returns the DAG in host memory (FCpuDag, every
    # level copied out, src/DAGCompression.cu:1073-1074, 1089): create_dag +
    # svdag_dag_to_host, from device-resident keys.
    dropin_times = []
    for r in range(4):
        _lib.hip_check(H.hipDeviceSynchronize(), "hipDeviceSynchronize")
        t = time.perf_counter()
        dag = _lib.SvdagDag()
        _lib.check(L.svdag_create_dag(ctx.handle, d_keys.ptr, n, depth, ctypes.byref(dag)))
        host = _lib.SvdagDag()
        _lib.check(L.svdag_dag_to_host(ctx.handle, ctypes.byref(dag), ctypes.byref(host)))
        dt = time.perf_counter() - t
        if r:
            dropin_times.append(dt)
        L.svdag_free_dag(ctx.handle, ctypes.byref(host))
        L.svdag_free_dag(ctx.handle, ctypes.byref(dag))
    dropin = {"ms": round(1e3 * float(np.median(dropin_times)), 3),
              "voxels_per_s": round(n / float(np.median(dropin_times)), 1),
              "what": "create_dag + svdag_dag_to_host (the level arrays in host memory, as CreateDAG returns "
                      "them), keys device-resident; median of 3 after 1 warm-up"}
    log(f"drop-in (create_dag + copy to host): {dropin['ms']:.2f} ms")

    vox_stats = None if args.no_voxelize else bench_voxelize(svd, spec, args, ctx)
    trace_stats = None if args.no_trace else bench_trace(svd, result_words, depth, args, ctx)
    sort_stats = None if args.no_sort else bench_sort(svd, args, ctx)

    chunk_stats = None
    if not args.no_chunks:
        chunk_stats = bench_chunks(svd, _lib, args)
        if "MISMATCH" in chunk_stats["check"]:
            log(f"chunk workload: {chunk_stats['check']}")
    noise_stats = None
    if not args.no_noise:
        noise_stats = bench_noise(svd, _lib, args)
        if "MISMATCH" in noise_stats["check"]:
            log(f"noise workload: {noise_stats['check']}")

    out = {
        "metric": METRIC, "value": round(n / (ms_per_step / 1e3), 1), "unit": "voxels/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64", "data": "synthetic",
        "config": {"workload": spec["desc"], "n_keys": n, "depth": depth, "mode": "plain",
                   "parallelism": "single GPU"},
        "roofline": roofline, "cpu_baseline": cpu, "check": check, "chunks": chunk_stats,
        "dropin": dropin, "noise": noise_stats, "voxelize": vox_stats, "trace": trace_stats,
        "sort_keys": sort_stats,
    }
    print(json.dumps(out), flush=True)
    ctx.close()
    ok = "MISMATCH" not in check and all(x is None or "MISMATCH" not in x["check"] for x in (chunk_stats, noise_stats))
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
