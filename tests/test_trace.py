"""Primary rays through the flat layout (the reference's trace_paths,
src/DAGTracer.cu:177-428, camera of src/DAGTracer.h:17-82).

The oracle's walk is pinned by what geometry fixes, independently of its own
arithmetic: every hit is a voxel of the decoded set, it is the voxel the ray
enters first (brute force over every voxel, up to float32 rounding and the
walk's own 1e-4 plane tolerance), and a ray that clearly misses every voxel
reports no hit.  A planted child-pointer swap in the words fails those checks.
The device kernel must then equal the oracle pixel for pixel (integer output:
tolerance 0) on the CPU emulator and on the GPU, including the colours in both
of the reference's modes and a layout whose walk leaves the words.
"""
import numpy as np
import pytest

from dagcheck import decode_words, morton

ALL_ONES = 0xFFFFFFFF


def scene_keys(depth, seed):
    """Solid balls and a slab in a 2^depth grid (occluders for each other)."""
    side = 1 << depth
    rng = np.random.default_rng(seed)
    g = np.arange(side, dtype=np.float64) + 0.5
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    occ = np.zeros((side,) * 3, dtype=bool)
    for _ in range(5):
        c = rng.uniform(0.15 * side, 0.85 * side, 3)
        r = rng.uniform(0.06 * side, 0.2 * side)
        occ |= (x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2 < r * r
    occ |= np.abs(z - 0.3 * side) < 0.8  # a thin slab
    xi, yi, zi = np.nonzero(occ)
    return morton(xi.astype(np.uint64), yi.astype(np.uint64), zi.astype(np.uint64))


def camera(depth, eye, target, width, height, fov_deg=60.0, debug_level=0):
    """Pinhole camera in grid units, built as the reference's GetTraceParams
    with an identity box transform."""
    eye, target = np.asarray(eye, float), np.asarray(target, float)
    fwd = target - eye
    fwd /= np.linalg.norm(fwd)
    up0 = np.array([0.0, 0.0, 1.0]) if abs(fwd[2]) < 0.9 else np.array([0.0, 1.0, 0.0])
    right = np.cross(fwd, up0)
    right /= np.linalg.norm(right)
    up = np.cross(right, fwd)
    half = np.radians(fov_deg) / 2
    X = right * np.sin(half) * (width / height)
    Y = up * np.sin(half)
    Z = fwd * np.cos(half)
    bl = eye + Z - Y - X
    return {"debug_level": debug_level, "width": width, "height": height, "camera_position": tuple(eye),
            "ray_min": tuple(bl), "ray_ddx": tuple(2 * X / width), "ray_ddy": tuple(2 * Y / height)}


def ray_dirs(p):
    w, h = p["width"], p["height"]
    xs, ys = np.meshgrid(np.arange(w, dtype=np.float64), np.arange(h, dtype=np.float64))
    v = (np.asarray(p["ray_min"])[None, None] + xs[..., None] * np.asarray(p["ray_ddx"])
         + ys[..., None] * np.asarray(p["ray_ddy"]) - np.asarray(p["camera_position"]))
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def box_spans(o, d, boxes):
    """Entry and exit distance (t >= 0) of rays o + t d through unit boxes."""
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = 1.0 / d
        t0 = (boxes[None, :, :] - o[None, None, :]) * inv[:, None, :]
        t1 = (boxes[None, :, :] + 1.0 - o[None, None, :]) * inv[:, None, :]
    lo = np.nanmax(np.minimum(t0, t1), axis=-1)
    hi = np.nanmin(np.maximum(t0, t1), axis=-1)
    return np.maximum(lo, 0.0), hi


def check_against_geometry(paths, voxels_xyz, p, tol=2e-3, max_grazing=0.002):
    """Every hit is a set voxel; it is entered first up to `tol`; clear misses
    report no hit.  Returns the number of grazing rays allowed through."""
    o = np.asarray(p["camera_position"], dtype=np.float64)
    dirs = ray_dirs(p).reshape(-1, 3)
    got = paths.reshape(-1, 3).astype(np.int64)
    vox = np.asarray(voxels_xyz, dtype=np.float64)
    index = {tuple(v): i for i, v in enumerate(np.asarray(voxels_xyz, dtype=np.int64))}
    hit = got.any(axis=1)
    # A hit at the origin voxel is indistinguishable from a miss; the scenes
    # keep (0, 0, 0) empty.
    assert (0, 0, 0) not in index
    bad = 0
    for s in range(0, len(dirs), 256):
        d = dirs[s:s + 256]
        enter, leave = box_spans(o, d, vox)
        crossing = leave - enter  # > 0: the ray passes through the box
        first = np.where(crossing > -tol, enter, np.inf).min(axis=1)
        clear = np.where(crossing > tol, enter, np.inf)
        for j in range(len(d)):
            r = s + j
            if hit[r]:
                key = tuple(got[r])
                assert key in index, f"ray {r}: hit {key} is not a voxel of the set"
                k = index[key]
                ok = crossing[j, k] > -tol and enter[j, k] <= first[j] + tol and enter[j, k] <= clear[j].min() + tol
            else:
                ok = not np.isfinite(clear[j]).any()
            bad += not ok
    assert bad <= max_grazing * len(dirs), f"{bad} of {len(dirs)} rays disagree with the geometry"
    return bad


def voxel_xyz(codes, depth):
    c = np.asarray(sorted(codes), dtype=np.uint64)
    xyz = np.zeros((c.size, 3), dtype=np.int64)
    for b in range(depth):
        for a, off in enumerate((2, 1, 0)):
            xyz[:, a] |= ((c >> np.uint64(3 * b + off)) & np.uint64(1)).astype(np.int64) << b
    return xyz


VIEWS = [
    # eye outside the volume looking in, eye inside, axis-aligned view, grazing the slab
    ((-20.0, -13.0, 41.0), (16.0, 16.0, 12.0)),
    ((15.3, 17.9, 20.2), (3.0, 30.0, 1.0)),
    ((16.25, 16.5, 70.0), (16.25, 16.5, 0.0)),
    ((60.0, 7.1, 9.7), (0.0, 21.0, 9.55)),
]


def test_oracle_walk_against_geometry(oracle):
    depth = 5
    keys = scene_keys(depth, 1)
    words = oracle.create_final_dag(oracle.create_dag(keys, depth))
    codes = decode_words(words, depth)
    assert codes == set(int(k) for k in np.unique(keys))
    xyz = voxel_xyz(codes, depth)
    for eye, target in VIEWS:
        p = camera(depth, eye, target, 64, 48)
        paths, _ = oracle.trace_paths(words, depth, p)
        assert paths.reshape(-1, 3).any(axis=1).sum() > 100
        check_against_geometry(paths, xyz, p)


def test_planted_pointer_swap_is_caught(oracle):
    depth = 5
    keys = scene_keys(depth, 1)
    words = oracle.create_final_dag(oracle.create_dag(keys, depth)).copy()
    xyz = voxel_xyz(decode_words(words, depth), depth)
    # Swap the root's first two child pointers: the walk then descends into
    # the wrong sub-volumes.
    mask = int(words[0])
    assert bin(mask).count("1") >= 2
    words[1], words[2] = words[2], words[1]
    p = camera(depth, *VIEWS[0], 64, 48)
    paths, _ = oracle.trace_paths(words, depth, p)
    with pytest.raises(AssertionError):
        check_against_geometry(paths, xyz, p)


def test_oracle_misses_and_colours(oracle):
    depth = 5
    keys = scene_keys(depth, 2)
    words = oracle.create_final_dag(oracle.create_dag(keys, depth))
    # Looking away from the volume: nothing is hit.
    p = camera(depth, (-10.0, -10.0, -10.0), (-30.0, -25.0, -40.0), 32, 16)
    paths, colors = oracle.trace_paths(words, depth, p)
    assert not paths.any()
    # Empty layout: every ray misses.
    paths, _ = oracle.trace_paths(np.zeros(0, dtype=np.uint32), depth, camera(depth, *VIEWS[0], 16, 16))
    assert not paths.any()
    # Position colours: each channel is the position's low byte up to the
    # rounding of x / 255 * 255.
    p = camera(depth, *VIEWS[0], 48, 32, debug_level=ALL_ONES)
    paths, colors = oracle.trace_paths(words, depth, p)
    for a in range(3):
        chan = (colors >> np.uint32(8 * a)) & np.uint32(0xFF)
        assert np.all(np.abs(chan.astype(np.int64) - (paths[..., a] & 0xFF).astype(np.int64)) <= 1)


def test_trace_params_from_view():
    import dag_compression_fast_amd as svd
    # Camera at the origin of a unit box looking along +y, 90 degree field of
    # view: the corner ray points along (-1, 1, -1) in a 2^depth grid.
    p = svd.trace_params((0.5, -1.0, 0.5), (1, 0, 0), (0, 0, 1), (0, 1, 0), 90.0, 100, 100,
                         (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 8)
    s = 256.0
    h = np.sin(np.pi / 4)
    assert np.allclose(p["camera_position"], (0.5 * s, -1.0 * s, 0.5 * s))
    assert np.allclose(np.subtract(p["ray_min"], p["camera_position"]), (-h * s, h * s, -h * s))
    assert np.allclose(np.multiply(p["ray_ddx"], 100), (2 * h * s, 0, 0))
    assert np.allclose(np.multiply(p["ray_ddy"], 100), (0, 0, 2 * h * s))
    from dag_compression_fast_amd import SvdagError
    with pytest.raises(SvdagError):
        svd.trace_params((0, 0, 0), (1, 0, 0), (0, 0, 1), (0, 1, 0), 90.0, 0, 10, (0, 0, 0), (1, 1, 1), 8)


def trace_cases(oracle, depths=(5, 7)):
    for depth in depths:
        keys = scene_keys(depth, depth)
        words = oracle.create_final_dag(oracle.create_dag(keys, depth))
        side = 1 << depth
        for i, (eye, target) in enumerate(VIEWS):
            sc = side / 32.0
            dbg = (0, ALL_ONES, 1)[i % 3]
            yield depth, words, camera(depth, np.multiply(eye, sc), np.multiply(target, sc), 40 + 8 * i, 24 + 4 * i,
                                       debug_level=dbg)


def check_device_equals_oracle(m, oracle, depths):
    for depth, words, p in trace_cases(oracle, depths):
        want_p, want_c = oracle.trace_paths(words, depth, p)
        got_p, got_c = m.trace_paths(words, depth, p)
        assert np.array_equal(got_p, want_p), f"depth {depth}: {np.sum(np.any(got_p != want_p, axis=-1))} pixels"
        assert np.array_equal(got_c, want_c), f"depth {depth} colours"
    # A child pointer past the end of the words: reported, not followed.
    depth, words, p = next(trace_cases(oracle, (5,)))
    bad = words.copy()
    bad[1] = np.uint32(bad.size + 5)
    want_p, want_c = oracle.trace_paths(bad, depth, p)
    got_p, got_c = m.trace_paths(bad, depth, p)
    assert (want_p == ALL_ONES).all(axis=-1).any()
    assert np.array_equal(got_p, want_p) and np.array_equal(got_c, want_c)
    empty_p, _ = m.trace_paths(np.zeros(0, dtype=np.uint32), depth, p)
    assert not empty_p.any()


def test_trace_on_the_cpu_emulator(oracle):
    import emu_ctypes
    m = emu_ctypes.use_emulator()
    try:
        check_device_equals_oracle(m, oracle, (5,))
    finally:
        from dag_compression_fast_amd import _lib
        import dag_compression_fast_amd as svd
        _lib._lib = None
        _lib._hip = None
        svd._default = None


@pytest.mark.gpu
def test_trace_on_gpu(gpu, oracle):
    print("[trace] small scenes", flush=True)
    check_device_equals_oracle(gpu, oracle, (5, 7, 9))
    # The 2048^3 triangle workload's DAG, 512 x 384 pixels from outside.
    print("[trace] triangle workload", flush=True)
    keys = gpu.gen_triangle_surfaces(0x5EED0001, 11, 20000, 4.0, 1.0, 0.3)
    words = gpu.create_final_dag(gpu.create_dag(keys, 11))
    for eye, target, dbg in (((-900.0, -700.0, 2600.0), (1024.0, 1024.0, 1024.0), ALL_ONES),
                             ((1000.5, 1100.25, 1050.75), (0.0, 2048.0, 300.0), 0)):
        p = camera(11, eye, target, 512, 384, debug_level=dbg)
        want_p, want_c = oracle.trace_paths(words, 11, p)
        got_p, got_c = gpu.trace_paths(words, 11, p)
        assert np.array_equal(got_p, want_p), f"{np.sum(np.any(got_p != want_p, axis=-1))} pixels differ"
        assert np.array_equal(got_c, want_c)
        assert want_p.reshape(-1, 3).any(axis=1).mean() > 0.2
