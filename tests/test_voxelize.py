"""Device fragment producer: triangle soup -> Morton keys.

The device voxeliser must emit exactly the voxel set of the host rule
(svdag_gen_voxelize, the rule behind the triangle-surface workload): the same
floating-point steps in the same order, no fused multiply-adds.  Checked as
sets (integer output: tolerance 0) on seeded soups, on hand-made edge cases
(axis-aligned, degenerate, outside or straddling the grid), and at the full
2048^3 workload on the GPU, where the DAG built from the device keys must equal
the one built from the host keys.  The host rule itself is pinned against its
geometric definition, exhaustively in exact arithmetic and by closed forms, in
tests/test_voxel_rule.py.
"""
import numpy as np
import pytest


def edge_cases(side):
    s = float(side)
    return np.array([
        # axis-aligned (normal along z, x, y): two projections edge-on
        [[1.0, 1.0, 2.5], [s - 2, 1.0, 2.5], [1.0, s - 2, 2.5]],
        [[3.25, 0.5, 0.5], [3.25, s - 1, 0.5], [3.25, 0.5, s - 1]],
        [[0.5, 7.0, 0.5], [s - 1, 7.0, 0.5], [0.5, 7.0, s - 1]],
        # corners on voxel boundaries
        [[0.0, 0.0, 0.0], [8.0, 0.0, 0.0], [0.0, 8.0, 8.0]],
        # degenerate: zero area (collinear), a point
        [[1.0, 1.0, 1.0], [2.0, 2.0, 2.0], [3.0, 3.0, 3.0]],
        [[4.0, 4.0, 4.0], [4.0, 4.0, 4.0], [4.0, 4.0, 4.0]],
        # straddling and outside the grid
        [[-5.0, 3.0, 2.0], [s + 4, 5.0, 9.0], [4.0, s + 3, -2.0]],
        [[s + 1, s + 2, s + 3], [s + 5, s + 2, s + 3], [s + 1, s + 7, s + 9]],
        # thin sliver
        [[0.3, 0.3, 0.3], [s - 0.7, 0.31, 0.35], [0.32, 0.33, s - 0.6]],
    ], dtype=np.float64)


def near_axis_soup(side, count, seed):
    """Triangles with an edge parallel to a grid axis up to 1e-13..1e-6: the
    projected edge normals then have a tiny component, where the per-row column
    ranges lean on their rounding margin."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        v0 = rng.uniform(1.0, side - 2.0, 3)
        ax = rng.integers(0, 3)
        v1 = v0.copy()
        v1[ax] = rng.uniform(1.0, side - 2.0)
        v1 += rng.choice([-1.0, 1.0], 3) * 10.0 ** rng.uniform(-13, -6, 3)
        v2 = v0 + rng.uniform(-0.3, 0.3, 3) * side
        if i % 3 == 0:  # also nearly parallel to the first edge
            v2 = v0 + (v1 - v0) * rng.uniform(0.2, 0.9) + 10.0 ** rng.uniform(-12, -2, 3)
        out.append([v0, v1, np.clip(v2, 0.0, side - 0.5)])
    return np.array(out, dtype=np.float64)


def _lib_opt():
    from dag_compression_fast_amd import _lib
    return _lib.OPT_VOXEL_STAGE_BYTES


def unique_keys(buf, n):
    keys = buf.to_numpy(np.uint64, n)
    buf.free()
    return np.unique(keys)


def test_voxelize_on_the_cpu_emulator():
    import dag_compression_fast_amd as host
    hostset = {}
    rng_cases = []
    for seed, depth, count, mu, sigma in ((0x5EED0101, 7, 120, 2.0, 0.7), (0x5EED0102, 9, 40, 3.0, 0.5)):
        tris = host.gen_triangles(seed, depth, count, mu, sigma)
        rng_cases.append((tris, depth))
        # The triangle-surface workload is this rule over these triangles.
        assert np.array_equal(host.gen_voxelize(tris, depth),
                              np.unique(host.gen_triangle_surfaces(seed, depth, count, mu, sigma, 0.0)))
    rng_cases.append((edge_cases(32), 5))
    rng_cases.append((near_axis_soup(128, 150, 3), 7))
    for tris, depth in rng_cases:
        hostset[id(tris)] = host.gen_voxelize(tris, depth)
    import emu_ctypes
    m = emu_ctypes.use_emulator()
    try:
        for tris, depth in rng_cases:
            buf, n = m.voxelize_triangles(tris, depth)
            got = unique_keys(buf, n)
            assert np.array_equal(got, hostset[id(tris)]), f"depth {depth}: {got.size} vs {hostset[id(tris)].size}"
            assert got.size == 0 or int(got.max()) < 8 ** depth
        buf, n = m.voxelize_triangles(np.zeros((0, 3, 3)), 5)
        assert n == 0
        bad = edge_cases(32)
        bad[2, 1, 0] = np.nan
        with pytest.raises(m.SvdagError):
            m.voxelize_triangles(bad, 5)
        # Two passes over the columns (no staging area): the same voxels, and
        # the same key sequence as the one-pass path.
        tris, depth = rng_cases[0]
        buf, n = m.voxelize_triangles(tris, depth)
        one_pass = buf.to_numpy(np.uint64, n)
        buf.free()
        ctx = m.default_context()
        ctx.set_option(_lib_opt(), 0)
        try:
            buf, n2 = m.voxelize_triangles(tris, depth)
            two_pass = buf.to_numpy(np.uint64, n2)
            buf.free()
        finally:
            ctx.set_option(_lib_opt(), 4 << 30)
        assert np.array_equal(one_pass, two_pass)
    finally:
        from dag_compression_fast_amd import _lib
        import dag_compression_fast_amd as svd
        _lib._lib = None
        _lib._hip = None
        svd._default = None


@pytest.mark.gpu
def test_voxelize_full_workload_on_gpu(gpu):
    # The 2048^3 triangle workload (20k triangles, edge lognormal(4, 1)).
    seed, depth, count, mu, sigma = 0x5EED0001, 11, 20000, 4.0, 1.0
    tris = gpu.gen_triangles(seed, depth, count, mu, sigma)
    print(f"[voxelize] {count} triangles depth {depth}", flush=True)
    want = gpu.gen_voxelize(tris, depth, threads=16)
    buf, n = gpu.voxelize_triangles(tris, depth)
    print(f"[voxelize] device keys {n}, host unique {want.size}", flush=True)
    try:
        got = np.unique(buf.to_numpy(np.uint64, n))
        assert np.array_equal(got, want)
        # The DAG of the device keys (with their duplicates) equals the DAG of
        # the host's unique keys.
        a = gpu.create_final_dag(gpu.create_dag(buf, depth))
    finally:
        buf.free()
    b = gpu.create_final_dag(gpu.create_dag(want, depth))
    assert np.array_equal(a, b)
    for tris_e, d in ((edge_cases(64), 6), (edge_cases(2048), 11), (near_axis_soup(2048, 400, 4), 11)):
        buf, n = gpu.voxelize_triangles(tris_e, d)
        assert np.array_equal(unique_keys(buf, n), gpu.gen_voxelize(tris_e, d))
    # The two-pass path (no staging area) gives the same key sequence.
    tris_e = near_axis_soup(2048, 400, 4)
    buf, n = gpu.voxelize_triangles(tris_e, 11)
    one_pass = buf.to_numpy(np.uint64, n)
    buf.free()
    ctx = gpu.default_context()
    ctx.set_option(_lib_opt(), 0)
    try:
        buf, n2 = gpu.voxelize_triangles(tris_e, 11)
        two_pass = buf.to_numpy(np.uint64, n2)
        buf.free()
    finally:
        ctx.set_option(_lib_opt(), 4 << 30)
    assert np.array_equal(one_pass, two_pass)
