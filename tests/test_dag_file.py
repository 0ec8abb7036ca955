"""The reference's on-disk DAG and point queries over the flat layout.

File framing: FFileWriter / FFileReader (src/Serializer.h:10-101); the fields
are read at src/DAG_Compression_Fast.cpp:362-375 in this order: six float64
(AABB min x, y, z, max x, y, z), uint32 level count, then the word array as
uint64 count + uint32 data.  The checks parse the bytes in that order, and
write files as those two readers lay them out, independently of the library.

Point queries walk the flat layout as the reference's tracer reads it
(src/DAGTracer.cu:85-176); they must agree with the decoded voxel set (tests/
dagcheck.py decode_words) on set and unset voxels, and report a walk that leaves
the words.
"""
import os
import struct

import numpy as np
import pytest

from dagcheck import decode_words

FIELDS = "<6dI"


def parse_reference_order(blob):
    """The file as src/DAG_Compression_Fast.cpp:362-375 reads it."""
    head = struct.calcsize(FIELDS)
    *bounds, levels = struct.unpack_from(FIELDS, blob, 0)
    (count,) = struct.unpack_from("<Q", blob, head)
    words = np.frombuffer(blob, dtype="<u4", count=count, offset=head + 8)
    assert len(blob) == head + 8 + 4 * count, "no bytes after the word array"
    return bounds[:3], bounds[3:], levels, words


def pack_reference_order(lo, hi, levels, words):
    w = np.asarray(words, dtype="<u4")
    return struct.pack(FIELDS, *lo, *hi, levels) + struct.pack("<Q", w.size) + w.tobytes()


@pytest.fixture(scope="module")
def svd():
    import dag_compression_fast_amd as m
    return m


def test_write_matches_reference_framing(svd, oracle, tmp_path):
    rng = np.random.default_rng(11)
    keys = rng.integers(0, 8 ** 6, size=4000, dtype=np.uint64)
    words = oracle.create_final_dag(oracle.create_dag(keys, 6))
    lo, hi = (-1.5, 0.25, 3.0), (10.0, 11.75, 14.5)
    path = tmp_path / "scene.dag.bin"
    svd.write_dag_file(path, words, 6, lo, hi)
    blob = path.read_bytes()
    assert blob == pack_reference_order(lo, hi, 6, words)
    got_lo, got_hi, levels, got_words = parse_reference_order(blob)
    assert tuple(got_lo) == lo and tuple(got_hi) == hi and levels == 6
    assert np.array_equal(got_words, words)


def test_read_round_trip_and_foreign_file(svd, tmp_path):
    rng = np.random.default_rng(12)
    words = rng.integers(0, 2 ** 32, size=1000, dtype=np.uint64).astype(np.uint32)
    path = tmp_path / "foreign.dag.bin"
    path.write_bytes(pack_reference_order((0.0, 1.0, 2.0), (3.0, 4.0, 5.0), 16, words))
    f = svd.read_dag_file(path)
    assert f["levels"] == 16 and f["aabb_min"] == (0.0, 1.0, 2.0) and f["aabb_max"] == (3.0, 4.0, 5.0)
    assert np.array_equal(f["words"], words)
    out = tmp_path / "again.dag.bin"
    svd.write_dag_file(out, f["words"], f["levels"], f["aabb_min"], f["aabb_max"])
    assert out.read_bytes() == path.read_bytes()


def test_empty_dag_file(svd, tmp_path):
    path = tmp_path / "empty.dag.bin"
    svd.write_dag_file(path, np.zeros(0, dtype=np.uint32), 3)
    assert len(path.read_bytes()) == struct.calcsize(FIELDS) + 8
    f = svd.read_dag_file(path)
    assert f["words"].size == 0 and f["levels"] == 3


def test_bad_files_are_errors(svd, tmp_path):
    from dag_compression_fast_amd import SvdagError
    with pytest.raises(SvdagError):
        svd.read_dag_file(tmp_path / "missing.dag.bin")
    path = tmp_path / "short.dag.bin"
    blob = pack_reference_order((0, 0, 0), (1, 1, 1), 5, np.arange(100, dtype=np.uint32))
    path.write_bytes(blob[:-4])  # one word short of its count
    with pytest.raises(SvdagError):
        svd.read_dag_file(path)
    path.write_bytes(blob[:20])  # inside the bounds
    with pytest.raises(SvdagError):
        svd.read_dag_file(path)


def check_queries(m, oracle, keys, depth, seed):
    """Queries of every input voxel, of random codes and of codes outside the
    volume agree with the decoded voxel set."""
    words = m.create_final_dag(m.create_dag(keys, depth))
    assert np.array_equal(words, oracle.create_final_dag(oracle.create_dag(keys, depth)))
    voxels = np.array(sorted(decode_words(words, depth)), dtype=np.uint64) if depth <= 8 else np.unique(keys)
    rng = np.random.default_rng(seed)
    probe = rng.integers(0, 8 ** depth, size=5000, dtype=np.uint64)
    codes = np.concatenate([keys, probe, np.array([8 ** depth, 2 ** 63], dtype=np.uint64)])
    got = m.query_voxels(words, depth, codes)
    want = np.isin(codes, voxels).astype(np.uint8)
    assert np.array_equal(got, want)
    assert got[:keys.size].all()
    return words


def test_query_on_the_cpu_emulator(oracle):
    import emu_ctypes
    m = emu_ctypes.use_emulator()
    try:
        rng = np.random.default_rng(13)
        for depth, n in ((3, 40), (5, 700), (8, 3000)):
            keys = rng.integers(0, 8 ** depth, size=n, dtype=np.uint64)
            words = check_queries(m, oracle, keys, depth, depth)
        # A pointer past the end: the walk reports it.
        bad = words.copy()
        bad[1] = np.uint32(bad.size + 7)
        first = np.array([0], dtype=np.uint64)
        slot0 = int(words[0]) & -int(words[0])  # lowest set slot of the root
        code = np.uint64((slot0.bit_length() - 1) << (3 * 7))
        assert m.query_voxels(bad, 8, np.array([code], dtype=np.uint64))[0] == 2
        assert m.query_voxels(np.zeros(0, dtype=np.uint32), 8, first)[0] == 0
    finally:
        from dag_compression_fast_amd import _lib
        import dag_compression_fast_amd as svd
        _lib._lib = None
        _lib._hip = None
        svd._default = None


@pytest.mark.gpu
def test_query_and_file_on_gpu(gpu, oracle, tmp_path):
    import ctypes
    from dag_compression_fast_amd import _lib
    rng = np.random.default_rng(14)
    for depth, n in ((6, 20000), (10, 200000)):
        print(f"[query] depth={depth} n={n}", flush=True)
        keys = rng.integers(0, 8 ** depth, size=n, dtype=np.uint64)
        keys = np.concatenate([keys, keys[: n // 5]])
        words = check_queries(gpu, oracle, keys, depth, 100 + depth)
    # Device-resident words written straight from HBM.
    print("[file] device words", flush=True)
    dw = gpu.DeviceBuffer.from_numpy(words)
    try:
        lo = (ctypes.c_double * 3)(0.0, 0.0, 0.0)
        hi = (ctypes.c_double * 3)(1.0, 1.0, 1.0)
        path = tmp_path / "gpu.dag.bin"
        _lib.check(_lib.lib().svdag_write_dag_file(gpu.default_context().handle, os.fsencode(path), lo, hi, 10,
                                                   dw.ptr, words.size, _lib.SVDAG_DEVICE))
    finally:
        dw.free()
    assert path.read_bytes() == pack_reference_order((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 10, words)
