"""MI355X (gfx950) sparse-voxel-DAG build: Python mirror of the reference API.

The reference exposes (Phyronnaz/DAG_Compression_Fast, src/DAGCompression.h:42-50)::

    FCpuDag CreateDAG(const TStaticArray<FMortonCode, GPU>& Fragments, int32 Depth);
    FCpuDag MergeDAGs(std::vector<FCpuDag>&& CpuDags);
    TStaticArray<uint32, GPU> CreateFinalDAG(FCpuDag&& CpuDag);
    void FreeAll();

Here the same four operations are ``create_dag``, ``merge_dags``,
``create_final_dag`` and ``free_all``, with the same argument meaning, running
the hand-written HIP kernels of ``libsvdag.so`` through its C ABI.  Errors raise
``SvdagError`` (the reference aborts).  Arrays are numpy on the host side;
device memory is held in ``DeviceBuffer`` objects.
"""
import ctypes
import os

import numpy as np

from . import _lib
from ._lib import SvdagError, SVDAG_DEVICE, SVDAG_HOST

__all__ = [
    "BatchedGraph", "Context", "DeviceBuffer", "DeviceView", "Dag", "ResidentDag", "SvdagError", "create_dag", "create_dag_batched",
    "create_dag_batched_gather", "merge_dags",
    "create_final_dag", "free_all", "default_context", "morton_encode", "gen_sphere_shells",
    "gen_triangle_surfaces", "gen_value_noise", "gen_chunk_plan", "gen_chunk",
]


class DeviceBuffer:
    """A hipMalloc'd device allocation."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = ctypes.c_void_p()
        _lib.hip_check(_lib.hip().hipMalloc(ctypes.byref(self.ptr), max(1, self.nbytes)), "hipMalloc")

    @classmethod
    def from_numpy(cls, arr):
        arr = np.ascontiguousarray(arr)
        buf = cls(arr.nbytes)
        buf.dtype = arr.dtype
        buf.count = arr.size
        if arr.nbytes:
            _lib.hip_check(_lib.hip().hipMemcpy(buf.ptr, arr.ctypes.data, arr.nbytes, _lib.HIP_H2D), "hipMemcpy")
        return buf

    @property
    def address(self):
        return self.ptr.value or 0

    def to_numpy(self, dtype, count, offset_bytes=0):
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            _lib.hip_check(_lib.hip().hipMemcpy(out.ctypes.data, self.address + offset_bytes, out.nbytes,
                                                _lib.HIP_D2H), "hipMemcpy")
        return out

    def free(self):
        if self.ptr is not None and self.ptr.value:
            _lib.hip().hipFree(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceView(DeviceBuffer):
    """A range of another DeviceBuffer (count elements of dtype from byte
    offset_bytes); it owns nothing, and freeing it does nothing."""

    def __init__(self, base, offset_bytes, count, dtype=np.uint64):
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        self.nbytes = self.count * self.dtype.itemsize
        if offset_bytes < 0 or offset_bytes + self.nbytes > base.nbytes:
            raise ValueError("view outside its buffer")
        self.ptr = ctypes.c_void_p(base.address + int(offset_bytes))
        self.base = base

    def free(self):
        self.ptr = None

    def __del__(self):
        pass


def _copy_d2h(addr, dtype, count):
    out = np.empty(count, dtype=dtype)
    if out.nbytes:
        _lib.hip_check(_lib.hip().hipMemcpy(out.ctypes.data, addr, out.nbytes, _lib.HIP_D2H), "hipMemcpy")
    return out


class Context:
    """A build context: stream binding, scratch arena (the reference's FPool)."""

    def __init__(self, stream=None):
        self.handle = ctypes.c_void_p()
        _lib.check(_lib.lib().svdag_context_create(ctypes.byref(self.handle), stream))

    def set_option(self, option, value):
        _lib.check(_lib.lib().svdag_set_option(self.handle, option, int(value)))

    def memory_usage(self):
        """Device bytes held: {"scratch", "results" (owned by live DAGs),
        "cached" (freed result blocks kept for reuse)}."""
        v = [ctypes.c_uint64() for _ in range(3)]
        _lib.check(_lib.lib().svdag_memory_usage(self.handle, *(ctypes.byref(x) for x in v)))
        return {"scratch": v[0].value, "results": v[1].value, "cached": v[2].value}

    def free_all(self):
        _lib.check(_lib.lib().svdag_free_all(self.handle))

    # Per-kernel timing (svdag_profile_*): HIP event pairs on the context's stream.
    @staticmethod
    def kernel_names():
        names = []
        while True:
            n = _lib.lib().svdag_kernel_name(len(names))
            if n is None:
                return names
            names.append(n.decode())

    def profile_enable(self, kernels=None):
        """Times every launch of the named kernels (None: all, []: off)."""
        names = self.kernel_names()
        sel = names if kernels is None else kernels
        mask = 0
        for k in sel:
            mask |= 1 << names.index(k)
        _lib.check(_lib.lib().svdag_profile_enable(self.handle, mask))

    def profile_reset(self):
        _lib.check(_lib.lib().svdag_profile_reset(self.handle))

    def profile_read(self):
        """{kernel: (total_ms, launches)} for the kernels timed since the last reset."""
        out = {}
        ms, cnt = ctypes.c_double(), ctypes.c_uint64()
        for i, name in enumerate(self.kernel_names()):
            _lib.check(_lib.lib().svdag_profile_read(self.handle, i, ctypes.byref(ms), ctypes.byref(cnt)))
            if cnt.value:
                out[name] = (ms.value, int(cnt.value))
        return out

    def close(self):
        if self.handle is not None and self.handle.value:
            _lib.lib().svdag_context_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default = None


def default_context():
    global _default
    if _default is None:
        _default = Context()
    return _default


class Dag:
    """A device- or host-resident DAG (mirror of FCpuDag) owned by Python."""

    def __init__(self, raw, ctx):
        self.raw = raw
        self.ctx = ctx

    @property
    def num_levels(self):
        return self.raw.num_levels

    @property
    def level_counts(self):
        return [int(self.raw.levels[i].num_nodes) for i in range(self.raw.num_levels)]

    @property
    def num_leaves(self):
        return int(self.raw.num_leaves)

    @property
    def pre_dedup(self):
        return [int(self.raw.pre_dedup[i]) for i in range(self.raw.num_levels)]

    @property
    def leaf_positions(self):
        return int(self.raw.leaf_positions)

    @property
    def hash_collisions(self):
        return int(self.raw.hash_collisions)

    @property
    def total_words(self):
        return 2 * self.num_leaves + sum(int(self.raw.level_words[i]) for i in range(self.raw.num_levels))

    def to_numpy(self):
        """Level arrays as numpy (the FCpuDag contents), same layout as the oracle's dict."""
        r = self.raw
        if r.location == SVDAG_DEVICE:
            get = _copy_d2h
        else:
            def get(addr, dtype, count):
                if count == 0:
                    return np.zeros(0, dtype=dtype)
                return np.ctypeslib.as_array(ctypes.cast(addr, ctypes.POINTER(np.ctypeslib.as_ctypes_type(dtype))),
                                             shape=(count,)).copy()
        levels = []
        for i in range(r.num_levels):
            lv = r.levels[i]
            n = int(lv.num_nodes)
            levels.append({
                "masks": get(lv.masks, np.uint8, n),
                "children": get(lv.children, np.uint32, 8 * n).reshape(n, 8),
                "hashes": get(lv.hashes, np.uint64, n),
            })
        return {"levels": levels, "leaves": get(r.leaves, np.uint64, int(r.num_leaves)),
                "hash_collisions": int(r.hash_collisions)}

    def to_host(self):
        """FCpuDag-like host copy (svdag_dag_to_host)."""
        out = _lib.SvdagDag()
        _lib.check(_lib.lib().svdag_dag_to_host(self.ctx.handle, ctypes.byref(self.raw), ctypes.byref(out)))
        return Dag(out, self.ctx)

    def to_device(self):
        out = _lib.SvdagDag()
        _lib.check(_lib.lib().svdag_dag_to_device(self.ctx.handle, ctypes.byref(self.raw), ctypes.byref(out)))
        return Dag(out, self.ctx)

    def free(self):
        if self.raw is not None and self.ctx.handle is not None:
            _lib.lib().svdag_free_dag(self.ctx.handle, ctypes.byref(self.raw))
        self.raw = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _as_device_keys(fragments):
    if isinstance(fragments, DeviceBuffer):
        return fragments, int(fragments.count), fragments.dtype
    arr = np.ascontiguousarray(fragments)
    if arr.dtype not in (np.uint32, np.uint64):
        arr = arr.astype(np.uint64)
    return DeviceBuffer.from_numpy(arr), arr.size, arr.dtype


def create_dag(fragments, depth, ctx=None):
    """DAGCompression::CreateDAG (src/DAGCompression.cu:1003-1107).

    fragments: unsorted Morton codes with duplicates, a numpy uint64/uint32 array
    or a DeviceBuffer (with ``dtype`` and ``count``).  Returns a device-resident Dag.
    """
    ctx = ctx or default_context()
    buf, n, dtype = _as_device_keys(fragments)
    out = _lib.SvdagDag()
    fn = _lib.lib().svdag_create_dag_u32 if dtype == np.uint32 else _lib.lib().svdag_create_dag
    _lib.check(fn(ctx.handle, buf.ptr, n, int(depth), ctypes.byref(out)))
    return Dag(out, ctx)


def create_dag_batched(keys, offsets, chunk_depth, ctx=None):
    """MergeDAGs(CreateDAG(chunk_i)) for chunks back to back in ``keys`` (chunk-local
    codes), chunk i at keys[offsets[i]:offsets[i+1]], chunks in Morton order."""
    ctx = ctx or default_context()
    buf, n, dtype = _as_device_keys(keys)
    if dtype != np.uint64:
        raise SvdagError(1, "batched keys must be uint64")
    offs = np.ascontiguousarray(offsets, dtype=np.uint64)
    out = _lib.SvdagDag()
    _lib.check(_lib.lib().svdag_create_dag_batched(
        ctx.handle, buf.ptr, offs.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), len(offs) - 1, int(chunk_depth),
        ctypes.byref(out)))
    return Dag(out, ctx)


def create_dag_batched_gather(chunk_keys, chunk_depth, ctx=None, num_chunks=None):
    """The batched build with each chunk's keys in its own DeviceBuffer, chunks
    in Morton order: a list (None for an empty chunk) or a dict {chunk index:
    DeviceBuffer} with num_chunks chunks in all."""
    ctx = ctx or default_context()
    items = chunk_keys.items() if isinstance(chunk_keys, dict) else enumerate(chunk_keys)
    num = int(num_chunks) if num_chunks is not None else len(chunk_keys)
    ptrs = np.zeros(num, dtype=np.uint64)
    counts = np.zeros(num, dtype=np.uint64)
    for i, b in items:
        if b is not None and b.count:
            if np.dtype(b.dtype) != np.uint64:
                raise SvdagError(1, "batched keys must be uint64")
            ptrs[i] = b.address
            counts[i] = b.count
    out = _lib.SvdagDag()
    _lib.check(_lib.lib().svdag_create_dag_batched_gather(
        ctx.handle, ptrs.ctypes.data_as(ctypes.POINTER(ctypes.c_void_p)),
        counts.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), num, int(chunk_depth), ctypes.byref(out)))
    return Dag(out, ctx)


class BatchedGraph:
    """The batched build captured once as a HIP graph (svdag_batched_graph_*)
    over the device buffers `keys` (capacity max_keys keys) and `offsets`
    (num_chunks + 1 uint64).  Write new chunk keys and offsets into those
    buffers, then ``launch()`` and ``result()``; the result's arrays live in the
    graph until its next launch."""

    def __init__(self, keys, offsets, num_chunks, chunk_depth, max_keys, ctx=None):
        self.ctx = ctx or default_context()
        self.handle = ctypes.c_void_p()
        _lib.check(_lib.lib().svdag_batched_graph_create(self.ctx.handle, keys.ptr, offsets.ptr, int(num_chunks),
                                                         int(chunk_depth), int(max_keys), ctypes.byref(self.handle)))

    def launch(self):
        _lib.check(_lib.lib().svdag_batched_graph_launch(self.handle))

    def result(self):
        out = _lib.SvdagDag()
        _lib.check(_lib.lib().svdag_batched_graph_result(self.handle, ctypes.byref(out)))
        return Dag(out, self.ctx)

    def close(self):
        if self.handle is not None and self.handle.value:
            _lib.lib().svdag_batched_graph_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 -- interpreter shutdown
            pass


class ResidentDag:
    """A global DAG kept on the device while chunks are added to it
    (svdag_resident_*): 8^slot_levels chunk slots of side 2^chunk_depth, one
    persistent dedup table per level, each for min(max_nodes, 8^l) distinct
    nodes.  ``insert`` adds the chunks of a round
    (a batched build over the slots) and returns each added chunk's root hash
    by slot; ``dag`` returns the global DAG in the reference's form (levels in
    ascending hash order), equal to MergeDAGs over the added chunks."""

    def __init__(self, chunk_depth, slot_levels, max_nodes, ctx=None):
        self.ctx = ctx or default_context()
        self.chunk_depth = int(chunk_depth)
        self.slot_levels = int(slot_levels)
        self.num_slots = 8 ** self.slot_levels
        self.handle = ctypes.c_void_p()
        _lib.check(_lib.lib().svdag_resident_create(self.ctx.handle, self.chunk_depth, self.slot_levels,
                                                    int(max_nodes), ctypes.byref(self.handle)))

    def insert(self, round_dag):
        hashes = np.zeros(self.num_slots, dtype=np.uint64)
        _lib.check(_lib.lib().svdag_resident_insert(self.handle, ctypes.byref(round_dag.raw),
                                                    hashes.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))))
        return hashes

    def counts(self):
        nodes = np.zeros(self.chunk_depth - 1, dtype=np.uint64)
        chunks = ctypes.c_uint64()
        _lib.check(_lib.lib().svdag_resident_counts(self.handle, nodes.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                                    ctypes.byref(chunks)))
        return nodes, int(chunks.value)

    def dag(self):
        out = _lib.SvdagDag()
        _lib.check(_lib.lib().svdag_resident_dag(self.handle, ctypes.byref(out)))
        return Dag(out, self.ctx)

    def close(self):
        if self.handle is not None and self.handle.value:
            _lib.lib().svdag_resident_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 -- interpreter shutdown
            pass


def merge_dags(dags, ctx=None):
    """DAGCompression::MergeDAGs: 8^k DAGs of equal depth in Morton order -> one Dag."""
    ctx = ctx or (dags[0].ctx if dags else default_context())
    arr = (_lib.SvdagDag * len(dags))()
    for i, d in enumerate(dags):
        arr[i] = d.raw
    out = _lib.SvdagDag()
    _lib.check(_lib.lib().svdag_merge_dags(ctx.handle, arr, len(dags), ctypes.byref(out)))
    return Dag(out, ctx)


def create_final_dag(dag, ctx=None):
    """DAGCompression::CreateFinalDAG -> numpy uint32 words of the flat layout."""
    ctx = ctx or dag.ctx
    d = dag if dag.raw.location == SVDAG_DEVICE else dag.to_device()
    words = ctypes.c_void_p()
    n = ctypes.c_uint64()
    _lib.check(_lib.lib().svdag_create_final_dag(ctx.handle, ctypes.byref(d.raw), ctypes.byref(words),
                                                 ctypes.byref(n)))
    try:
        return _copy_d2h(words.value or 0, np.uint32, int(n.value))
    finally:
        _lib.lib().svdag_free_words(ctx.handle, words)


def free_all(ctx=None):
    """DAGCompression::FreeAll: release the context's scratch arena."""
    (ctx or default_context()).free_all()


# --- the DAG file and point queries ------------------------------------------
def write_dag_file(path, words, levels, aabb_min=(0.0, 0.0, 0.0), aabb_max=(1.0, 1.0, 1.0)):
    """Writes the flat layout `words` (host uint32 array) as the reference's DAG
    file (FFileWriter framing, src/Serializer.h:10-101; fields in the order
    src/DAG_Compression_Fast.cpp:362-375 reads them)."""
    w = np.ascontiguousarray(words, dtype=np.uint32)
    lo = (ctypes.c_double * 3)(*aabb_min)
    hi = (ctypes.c_double * 3)(*aabb_max)
    _lib.check(_lib.lib().svdag_write_dag_file(None, os.fsencode(path), lo, hi, int(levels),
                                               w.ctypes.data if w.size else None, w.size, SVDAG_HOST))


def read_dag_file(path):
    """Reads a DAG file: dict(words, levels, aabb_min, aabb_max)."""
    lo = (ctypes.c_double * 3)()
    hi = (ctypes.c_double * 3)()
    levels = ctypes.c_uint32()
    p = ctypes.c_void_p()
    n = ctypes.c_uint64()
    _lib.check(_lib.lib().svdag_read_dag_file(os.fsencode(path), lo, hi, ctypes.byref(levels), ctypes.byref(p),
                                              ctypes.byref(n)))
    try:
        words = (np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ctypes.c_uint32)), shape=(n.value,)).copy()
                 if n.value else np.zeros(0, dtype=np.uint32))
    finally:
        _lib.lib().svdag_gen_free(p)
    return {"words": words, "levels": int(levels.value), "aabb_min": tuple(lo), "aabb_max": tuple(hi)}


def gen_triangles(seed, depth, count, mu, sigma):
    """The triangle soup of gen_triangle_surfaces: float64 array (count, 3, 3)."""
    out = np.zeros((max(0, int(count)), 3, 3), dtype=np.float64)
    _lib.check(_lib.lib().svdag_gen_triangles(seed, depth, int(count), mu, sigma,
                                              out.ctypes.data if out.size else None))
    return out


def gen_voxelize(triangles, depth, threads=8):
    """Host voxeliser (same rule as voxelize_triangles): sorted unique u64 keys."""
    t = np.ascontiguousarray(triangles, dtype=np.float64).reshape(-1, 9)
    p = ctypes.c_void_p()
    n = ctypes.c_uint64()
    rc = _lib.lib().svdag_gen_voxelize(t.ctypes.data if t.size else None, t.shape[0], int(depth), int(threads),
                                       ctypes.byref(p), ctypes.byref(n))
    return _take_keys(rc, p, n)


def voxelize_triangles(triangles, depth, ctx=None):
    """Device voxeliser: triangles (n, 3, 3) float64 in a 2^depth grid ->
    DeviceBuffer of unsorted u64 Morton keys with duplicates, and their count.
    The buffer feeds create_dag directly; free it with .free()."""
    ctx = ctx or default_context()
    t = np.ascontiguousarray(triangles, dtype=np.float64).reshape(-1, 9)
    if t.shape[0] == 0:
        return _KeyBuffer(ctypes.c_void_p(), 0, ctx), 0
    dt = DeviceBuffer.from_numpy(t)
    p = ctypes.c_void_p()
    n = ctypes.c_uint64()
    try:
        _lib.check(_lib.lib().svdag_voxelize_triangles(ctx.handle, dt.ptr, t.shape[0], int(depth), ctypes.byref(p),
                                                       ctypes.byref(n)))
    finally:
        dt.free()
    return _KeyBuffer(p, int(n.value), ctx), int(n.value)


class _KeyBuffer(DeviceBuffer):
    """Keys owned by the library (svdag_free_keys)."""

    def __init__(self, ptr, count, ctx):
        self.ptr = ptr
        self.nbytes = 8 * count
        self.count = count
        self.dtype = np.dtype(np.uint64)
        self._ctx = ctx

    def free(self):
        if self.ptr is not None and self.ptr.value:
            _lib.lib().svdag_free_keys(self._ctx.handle, self.ptr)
        self.ptr = None


def sort_keys(keys, bits, ctx=None):
    """The reference's SortKeys (src/DAGCompression.cu:336-357): uint64 or
    uint32 keys (numpy or DeviceBuffer) sorted stably on their low `bits` bits
    on the device.  Returns a numpy array for numpy input, else a new
    DeviceBuffer."""
    ctx = ctx or default_context()
    if not isinstance(keys, DeviceBuffer):
        keys = np.ascontiguousarray(keys)
    dt = np.dtype(keys.dtype)
    if dt not in (np.dtype(np.uint64), np.dtype(np.uint32)):
        raise SvdagError(1, "sort_keys takes uint64 or uint32 keys")
    buf, n, _ = _as_device_keys(keys)
    fn = _lib.lib().svdag_sort_keys if dt == np.uint64 else _lib.lib().svdag_sort_keys_u32
    out = None
    try:
        out = DeviceBuffer(dt.itemsize * n)
        out.dtype, out.count = dt, n
        _lib.check(fn(ctx.handle, buf.ptr, n, int(bits), out.ptr))
        if isinstance(keys, DeviceBuffer):
            result, out = out, None
            return result
        return out.to_numpy(dt, n)
    finally:
        if out is not None:
            out.free()
        if buf is not keys:
            buf.free()


def query_voxels(words, depth, codes, ctx=None):
    """Occupancy of Morton codes in a flat layout (host uint32 words of a
    depth-`depth` DAG), walked on the device as the reference's tracer reads
    it (src/DAGTracer.cu:85-176): uint8 per code, 1 set, 0 not set, 2 the walk
    left the words."""
    ctx = ctx or default_context()
    w = np.ascontiguousarray(words, dtype=np.uint32)
    c = np.ascontiguousarray(codes, dtype=np.uint64)
    if c.size == 0:
        return np.zeros(0, dtype=np.uint8)
    dw = DeviceBuffer.from_numpy(w) if w.size else None
    dc = DeviceBuffer.from_numpy(c)
    do = DeviceBuffer(c.size)
    try:
        _lib.check(_lib.lib().svdag_query_voxels(ctx.handle, dw.ptr if dw else None, w.size, int(depth), dc.ptr,
                                                 c.size, do.ptr))
        return do.to_numpy(np.uint8, c.size)
    finally:
        for b in (dw, dc, do):
            if b is not None:
                b.free()


def trace_params(position, right, up, direction, fov_degrees, width, height, aabb_min, aabb_max, depth,
                 debug_level=0):
    """The reference's FDAGTracer::GetTraceParams (src/DAGTracer.h:33-82):
    camera at `position` with rotation rows right, up, direction and full
    vertical field of view `fov_degrees`, in the 2^depth grid of the box
    [aabb_min, aabb_max].  Returns a dict for trace_paths."""
    def d3(v):
        return (ctypes.c_double * 3)(*[float(x) for x in v])
    out = _lib.TraceParams()
    _lib.check(_lib.lib().svdag_trace_params_from_view(d3(position), d3(right), d3(up), d3(direction),
                                                       float(fov_degrees), int(width), int(height), d3(aabb_min),
                                                       d3(aabb_max), int(depth), int(debug_level) & 0xFFFFFFFF,
                                                       ctypes.byref(out)))
    return out.as_dict()


def trace_paths(words, depth, params, ctx=None):
    """Primary rays through a flat layout on the device (the reference's
    trace_paths, src/DAGTracer.cu:177-428).  words: host uint32 words or a
    DeviceBuffer; params: dict from trace_params.  Returns (paths, colors):
    uint32 (height, width, 3) grid position of the first voxel hit ((0, 0, 0)
    for none, 0xFFFFFFFF when the walk leaves the words) and uint32
    (height, width) colours."""
    ctx = ctx or default_context()
    p = _lib.TraceParams.from_dict(params)
    own = not isinstance(words, DeviceBuffer)
    if own:
        w = np.ascontiguousarray(words, dtype=np.uint32)
        dw = DeviceBuffer.from_numpy(w) if w.size else None
        n_words = w.size
    else:
        dw, n_words = words, words.nbytes // 4
    pixels = int(p.width) * int(p.height)
    dpaths = DeviceBuffer(max(1, 12 * pixels))
    dcolors = DeviceBuffer(max(1, 4 * pixels))
    try:
        _lib.check(_lib.lib().svdag_trace_paths(ctx.handle, dw.ptr if dw else None, n_words, int(depth),
                                                ctypes.byref(p), dpaths.ptr, dcolors.ptr))
        paths = dpaths.to_numpy(np.uint32, 3 * pixels).reshape(p.height, p.width, 3)
        colors = dcolors.to_numpy(np.uint32, pixels).reshape(p.height, p.width)
        return paths, colors
    finally:
        for b in (dpaths, dcolors) + ((dw,) if own else ()):
            if b is not None:
                b.free()


# --- synthetic inputs --------------------------------------------------------
def morton_encode(x, y, z):
    return int(_lib.lib().svdag_morton_encode(x, y, z))


def _take_keys(rc, p, n):
    _lib.check(rc)
    try:
        if n.value == 0:
            return np.zeros(0, dtype=np.uint64)
        return np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ctypes.c_uint64)), shape=(n.value,)).copy()
    finally:
        _lib.lib().svdag_gen_free(p)


def gen_sphere_shells(seed, depth, count, rmin, rmax, dup_fraction, threads=8):
    p, n = ctypes.c_void_p(), ctypes.c_uint64()
    rc = _lib.lib().svdag_gen_sphere_shells(seed, depth, count, rmin, rmax, dup_fraction, threads, ctypes.byref(p),
                                            ctypes.byref(n))
    return _take_keys(rc, p, n)


def gen_triangle_surfaces(seed, depth, count, mu, sigma, dup_fraction, threads=8):
    p, n = ctypes.c_void_p(), ctypes.c_uint64()
    rc = _lib.lib().svdag_gen_triangle_surfaces(seed, depth, count, mu, sigma, dup_fraction, threads,
                                                ctypes.byref(p), ctypes.byref(n))
    return _take_keys(rc, p, n)


def gen_value_noise(seed, depth, spacing, sub_depth, sub_index, threads=8):
    p, n = ctypes.c_void_p(), ctypes.c_uint64()
    rc = _lib.lib().svdag_gen_value_noise(seed, depth, spacing, sub_depth, sub_index, threads, ctypes.byref(p),
                                          ctypes.byref(n))
    return _take_keys(rc, p, n)


def gen_chunk_plan(seed, num_chunks, median=2e5, sigma=1.0, nmin=1000, nmax=4_000_000):
    n = np.zeros(num_chunks, dtype=np.uint64)
    u = np.zeros(num_chunks, dtype=np.uint64)
    k = np.zeros(num_chunks, dtype=np.int32)
    _lib.check(_lib.lib().svdag_gen_chunk_plan(seed, num_chunks, median, sigma, nmin, nmax,
                                               n.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                               u.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                               k.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
    return n, u, k


def gen_chunk(seed, chunk, depth, median=2e5, sigma=1.0, nmin=1000, nmax=4_000_000, capacity=None):
    if capacity is None:
        capacity = int(gen_chunk_plan(seed, chunk + 1, median, sigma, nmin, nmax)[0][chunk])
    out = np.empty(capacity, dtype=np.uint64)
    n = ctypes.c_uint64()
    _lib.check(_lib.lib().svdag_gen_chunk(seed, chunk, depth, median, sigma, nmin, nmax, out.ctypes.data, capacity,
                                          ctypes.byref(n)))
    return out[: n.value]
