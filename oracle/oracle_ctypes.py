"""ctypes binding of the CPU oracle (oracle/svdag_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by the CPU-baseline leg of ``bench.py``, never by the product package.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libsvdag_oracle.so")


class OraLevel(ctypes.Structure):
    _fields_ = [
        ("masks", ctypes.POINTER(ctypes.c_uint8)),
        ("children", ctypes.POINTER(ctypes.c_uint32)),
        ("hashes", ctypes.POINTER(ctypes.c_uint64)),
        ("num_nodes", ctypes.c_uint64),
    ]


class OraDag(ctypes.Structure):
    _fields_ = [
        ("num_levels", ctypes.c_int32),
        ("levels", ctypes.POINTER(OraLevel)),
        ("leaves", ctypes.POINTER(ctypes.c_uint64)),
        ("num_leaves", ctypes.c_uint64),
        ("hash_collisions", ctypes.c_uint64),
        ("unique_keys", ctypes.c_uint64),
        ("leaf_positions", ctypes.c_uint64),
        ("pre_dedup", ctypes.c_uint64 * 32),
    ]


class OraTraceParams(ctypes.Structure):
    _fields_ = [
        ("debug_level", ctypes.c_uint32),
        ("width", ctypes.c_uint32),
        ("height", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
        ("camera_position", ctypes.c_double * 3),
        ("ray_min", ctypes.c_double * 3),
        ("ray_ddx", ctypes.c_double * 3),
        ("ray_ddy", ctypes.c_double * 3),
    ]


def build():
    """Compile the oracle with its Makefile (g++ only)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        u64p = ctypes.POINTER(ctypes.c_uint64)
        L.ora_murmur128xN.argtypes = [u64p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64, u64p]
        L.ora_murmur128xN.restype = None
        L.ora_hash_children.argtypes = [u64p]
        L.ora_hash_children.restype = ctypes.c_uint64
        L.ora_create_dag.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int32, ctypes.c_int32,
                                     ctypes.POINTER(OraDag)]
        L.ora_create_dag_u32.argtypes = L.ora_create_dag.argtypes
        L.ora_merge_dags.argtypes = [ctypes.POINTER(OraDag), ctypes.c_uint64, ctypes.c_int32,
                                     ctypes.POINTER(OraDag)]
        L.ora_create_final_dag.argtypes = [ctypes.POINTER(OraDag), ctypes.POINTER(ctypes.POINTER(ctypes.c_uint32)),
                                           u64p]
        L.ora_trace_paths.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int32,
                                      ctypes.POINTER(OraTraceParams), ctypes.c_int32, ctypes.c_void_p,
                                      ctypes.c_void_p]
        L.ora_free_dag.argtypes = [ctypes.POINTER(OraDag)]
        L.ora_free.argtypes = [ctypes.c_void_p]
        L.ora_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


class OracleError(RuntimeError):
    pass


def _check(rc):
    if rc != 0:
        raise OracleError(lib().ora_last_error().decode())


def _arr(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


def dag_to_py(d):
    """Copy an OraDag into a dict of numpy arrays."""
    levels = []
    for i in range(d.num_levels):
        lv = d.levels[i]
        n = int(lv.num_nodes)
        levels.append({
            "masks": _arr(lv.masks, n, np.uint8),
            "children": _arr(lv.children, 8 * n, np.uint32).reshape(n, 8),
            "hashes": _arr(lv.hashes, n, np.uint64),
        })
    return {
        "levels": levels,
        "leaves": _arr(d.leaves, int(d.num_leaves), np.uint64),
        "hash_collisions": int(d.hash_collisions),
        "unique_keys": int(d.unique_keys),
        "leaf_positions": int(d.leaf_positions),
        "pre_dedup": [int(d.pre_dedup[i]) for i in range(d.num_levels)],
    }


def py_to_dag(py, keep):
    """Build an OraDag view over a dict from dag_to_py; `keep` holds the buffers alive."""
    d = OraDag()
    n = len(py["levels"])
    lv_arr = (OraLevel * max(1, n))()
    for i, lv in enumerate(py["levels"]):
        m = np.ascontiguousarray(lv["masks"], dtype=np.uint8)
        c = np.ascontiguousarray(lv["children"], dtype=np.uint32).reshape(-1)
        h = np.ascontiguousarray(lv["hashes"], dtype=np.uint64)
        keep += [m, c, h]
        lv_arr[i].masks = m.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
        lv_arr[i].children = c.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
        lv_arr[i].hashes = h.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))
        lv_arr[i].num_nodes = len(h)
    leaves = np.ascontiguousarray(py["leaves"], dtype=np.uint64)
    keep += [lv_arr, leaves]
    d.num_levels = n
    d.levels = ctypes.cast(lv_arr, ctypes.POINTER(OraLevel))
    d.leaves = leaves.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))
    d.num_leaves = len(leaves)
    d.hash_collisions = int(py.get("hash_collisions", 0))
    return d


def create_dag(keys, depth, threads=1):
    """Oracle CreateDAG on a numpy uint64 / uint32 key array -> dict of numpy arrays."""
    keys = np.ascontiguousarray(keys)
    d = OraDag()
    fn = lib().ora_create_dag_u32 if keys.dtype == np.uint32 else lib().ora_create_dag
    if keys.dtype not in (np.uint32, np.uint64):
        keys = keys.astype(np.uint64)
        fn = lib().ora_create_dag
    _check(fn(keys.ctypes.data, len(keys), depth, threads, ctypes.byref(d)))
    try:
        return dag_to_py(d)
    finally:
        lib().ora_free_dag(ctypes.byref(d))


def merge_dags(dags, threads=1):
    """Oracle MergeDAGs over a list of dicts (Morton order, count a power of 8)."""
    keep = []
    arr = (OraDag * len(dags))()
    for i, py in enumerate(dags):
        arr[i] = py_to_dag(py, keep)
    out = OraDag()
    _check(lib().ora_merge_dags(arr, len(dags), threads, ctypes.byref(out)))
    try:
        return dag_to_py(out)
    finally:
        lib().ora_free_dag(ctypes.byref(out))


def create_final_dag(py):
    """Oracle CreateFinalDAG -> numpy uint32 word array."""
    keep = []
    d = py_to_dag(py, keep)
    words = ctypes.POINTER(ctypes.c_uint32)()
    n = ctypes.c_uint64()
    _check(lib().ora_create_final_dag(ctypes.byref(d), ctypes.byref(words), ctypes.byref(n)))
    try:
        return _arr(words, int(n.value), np.uint32)
    finally:
        lib().ora_free(words)


def murmur128xN(blocks_u64, seed, len_xor):
    data = np.ascontiguousarray(blocks_u64, dtype=np.uint64)
    out = (ctypes.c_uint64 * 2)()
    lib().ora_murmur128xN(data.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), len(data) // 2, seed, len_xor, out)
    return int(out[0]), int(out[1])


def hash_children(child_hashes):
    data = np.ascontiguousarray(child_hashes, dtype=np.uint64)
    assert data.shape == (8,)
    return int(lib().ora_hash_children(data.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))))


def trace_paths(words, depth, params, threads=8):
    """Oracle trace_paths: params is a dict with the fields of OraTraceParams.
    Returns (paths uint32 (height, width, 3), colors uint32 (height, width))."""
    w = np.ascontiguousarray(words, dtype=np.uint32)
    p = OraTraceParams()
    p.debug_level = int(params.get("debug_level", 0))
    p.width, p.height = int(params["width"]), int(params["height"])
    for f in ("camera_position", "ray_min", "ray_ddx", "ray_ddy"):
        getattr(p, f)[:] = [float(v) for v in params[f]]
    paths = np.zeros((p.height, p.width, 3), dtype=np.uint32)
    colors = np.zeros((p.height, p.width), dtype=np.uint32)
    rc = lib().ora_trace_paths(w.ctypes.data if w.size else None, w.size, depth, ctypes.byref(p), threads,
                               paths.ctypes.data, colors.ctypes.data)
    if rc != 0:
        raise OracleError("ora_trace_paths: bad arguments")
    return paths, colors
