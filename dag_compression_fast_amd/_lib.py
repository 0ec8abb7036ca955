"""ctypes binding of libsvdag.so (include/svdag.h) and of the HIP runtime calls
the Python layer needs for device buffers (hipMalloc / hipMemcpy / hipFree).

The library is loaded from the package's ``_build`` directory and nowhere else;
a missing library raises immediately (there is no fallback path).
"""
import ctypes
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
# SVDAG_LIBRARY selects another build of the library (default: the in-tree one).
LIB_PATH = os.environ.get("SVDAG_LIBRARY") or os.path.join(HERE, "_build", "libsvdag.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "svdag.h")

MAX_LEVELS = 24

SVDAG_HOST = 0
SVDAG_DEVICE = 1

OPT_CHECK_COLLISIONS = 1
OPT_MEMORY_BUDGET = 2
OPT_TABLE_LOAD_SHIFT = 3
OPT_HOST_THREADS = 4
OPT_WAIT_SECONDS = 5
OPT_RADIX_BITS = 6
OPT_SORT_TILE = 7
OPT_TABLE_ESTIMATE = 8
OPT_VOXEL_STAGE_BYTES = 9
OPT_SEG_FIRST_BITS = 10
OPT_TOP_DIGIT_FIRST = 11

STATUS_NAMES = {0: "OK", 1: "ARGUMENT", 2: "HIP", 3: "MEMORY", 4: "INTERNAL", 5: "CAPACITY"}


class SvdagLevel(ctypes.Structure):
    _fields_ = [
        ("masks", ctypes.c_void_p),
        ("children", ctypes.c_void_p),
        ("hashes", ctypes.c_void_p),
        ("num_nodes", ctypes.c_uint64),
    ]


class SvdagDag(ctypes.Structure):
    _fields_ = [
        ("num_levels", ctypes.c_int32),
        ("location", ctypes.c_int32),
        ("levels", SvdagLevel * MAX_LEVELS),
        ("leaves", ctypes.c_void_p),
        ("num_leaves", ctypes.c_uint64),
        ("level_words", ctypes.c_uint64 * MAX_LEVELS),
        ("hash_collisions", ctypes.c_uint64),
        ("num_keys", ctypes.c_uint64),
        ("leaf_positions", ctypes.c_uint64),
        ("pre_dedup", ctypes.c_uint64 * MAX_LEVELS),
        ("storage", ctypes.c_void_p),
    ]


class TraceParams(ctypes.Structure):
    """svdag_trace_params."""
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

    def as_dict(self):
        d = {"debug_level": self.debug_level, "width": self.width, "height": self.height}
        for f in ("camera_position", "ray_min", "ray_ddx", "ray_ddy"):
            d[f] = tuple(getattr(self, f))
        return d

    @classmethod
    def from_dict(cls, d):
        p = cls()
        p.debug_level = int(d.get("debug_level", 0))
        p.width, p.height = int(d["width"]), int(d["height"])
        for f in ("camera_position", "ray_min", "ray_ddx", "ray_ddy"):
            getattr(p, f)[:] = [float(v) for v in d[f]]
        return p


class SvdagError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"svdag error {STATUS_NAMES.get(code, code)}: {message}")
        self.code = code


_lib = None
_hip = None


def header_functions():
    """Names of every function declared in include/svdag.h."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(svdag_[a-z_0-9]+)\s*\(", text, re.M)))


def lib():
    """The loaded libsvdag.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run dag_compression_fast_amd.build.build() first")
        _lib = bind_library(ctypes.CDLL(LIB_PATH))
    return _lib


def bind_library(L):
    """Declares the C signatures of include/svdag.h on a loaded library."""
    vp, u64, i32, i64 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int32, ctypes.c_int64
    dagp = ctypes.POINTER(SvdagDag)
    dp = ctypes.POINTER(ctypes.c_double)
    sig = {
        "svdag_context_create": ([ctypes.POINTER(vp), vp], i32),
        "svdag_context_destroy": ([vp], i32),
        "svdag_set_option": ([vp, i32, i64], i32),
        "svdag_scratch_bytes": ([u64, i32, i32], u64),
        "svdag_create_dag": ([vp, vp, u64, i32, dagp], i32),
        "svdag_create_dag_u32": ([vp, vp, u64, i32, dagp], i32),
        "svdag_create_dag_noalloc": ([vp, vp, u64, i32, vp, u64, dagp], i32),
        "svdag_result_bytes": ([u64, i32], u64),
        "svdag_merge_dags": ([vp, dagp, u64, dagp], i32),
        "svdag_create_dag_batched": ([vp, vp, ctypes.POINTER(u64), u64, i32, dagp], i32),
        "svdag_create_dag_batched_gather": ([vp, ctypes.POINTER(vp), ctypes.POINTER(u64), u64, i32, dagp], i32),
        "svdag_batched_graph_create": ([vp, vp, vp, u64, i32, u64, ctypes.POINTER(vp)], i32),
        "svdag_batched_graph_launch": ([vp], i32),
        "svdag_batched_graph_result": ([vp, dagp], i32),
        "svdag_batched_graph_destroy": ([vp], i32),
        "svdag_resident_create": ([vp, i32, i32, u64, ctypes.POINTER(vp)], i32),
        "svdag_resident_destroy": ([vp], i32),
        "svdag_resident_insert": ([vp, dagp, ctypes.POINTER(u64)], i32),
        "svdag_resident_counts": ([vp, ctypes.POINTER(u64), ctypes.POINTER(u64)], i32),
        "svdag_resident_dag": ([vp, dagp], i32),
        "svdag_create_final_dag": ([vp, dagp, ctypes.POINTER(vp), ctypes.POINTER(u64)], i32),
        "svdag_dag_to_host": ([vp, dagp, dagp], i32),
        "svdag_dag_to_device": ([vp, dagp, dagp], i32),
        "svdag_free_dag": ([vp, dagp], i32),
        "svdag_free_words": ([vp, vp], i32),
        "svdag_free_all": ([vp], i32),
        "svdag_memory_usage": ([vp, ctypes.POINTER(u64), ctypes.POINTER(u64), ctypes.POINTER(u64)], i32),
        "svdag_last_error": ([], ctypes.c_char_p),
        "svdag_profile_enable": ([vp, u64], i32),
        "svdag_profile_reset": ([vp], i32),
        "svdag_profile_read": ([vp, i32, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(u64)], i32),
        "svdag_kernel_name": ([i32], ctypes.c_char_p),
        "svdag_morton_encode": ([ctypes.c_uint32] * 3, u64),
        "svdag_gen_sphere_shells": ([u64, i32, i32, ctypes.c_double, ctypes.c_double, ctypes.c_double, i32,
                                     ctypes.POINTER(vp), ctypes.POINTER(u64)], i32),
        "svdag_gen_triangle_surfaces": ([u64, i32, i32, ctypes.c_double, ctypes.c_double, ctypes.c_double, i32,
                                         ctypes.POINTER(vp), ctypes.POINTER(u64)], i32),
        "svdag_gen_value_noise": ([u64, i32, i32, ctypes.c_uint32, u64, i32, ctypes.POINTER(vp),
                                   ctypes.POINTER(u64)], i32),
        "svdag_gen_chunk_plan": ([u64, u64, ctypes.c_double, ctypes.c_double, u64, u64, ctypes.POINTER(u64),
                                  ctypes.POINTER(u64), ctypes.POINTER(i32)], i32),
        "svdag_gen_chunk": ([u64, u64, i32, ctypes.c_double, ctypes.c_double, u64, u64, vp, u64,
                             ctypes.POINTER(u64)], i32),
        "svdag_gen_free": ([vp], None),
        "svdag_write_dag_file": ([vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                  ctypes.c_uint32, vp, u64, i32], i32),
        "svdag_read_dag_file": ([ctypes.c_char_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                 ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(vp), ctypes.POINTER(u64)], i32),
        "svdag_query_voxels": ([vp, vp, u64, i32, vp, u64, vp], i32),
        "svdag_sort_keys": ([vp, vp, u64, i32, vp], i32),
        "svdag_sort_keys_u32": ([vp, vp, u64, i32, vp], i32),
        "svdag_voxelize_triangles": ([vp, vp, u64, i32, ctypes.POINTER(vp), ctypes.POINTER(u64)], i32),
        "svdag_free_keys": ([vp, vp], i32),
        "svdag_trace_params_from_view": ([dp, dp, dp, dp, ctypes.c_double, ctypes.c_uint32, ctypes.c_uint32, dp, dp,
                                          i32, ctypes.c_uint32, ctypes.POINTER(TraceParams)], i32),
        "svdag_trace_paths": ([vp, vp, u64, i32, ctypes.POINTER(TraceParams), vp, vp], i32),
        "svdag_gen_triangles": ([u64, i32, i32, ctypes.c_double, ctypes.c_double, vp], i32),
        "svdag_gen_voxelize": ([vp, u64, i32, i32, ctypes.POINTER(vp), ctypes.POINTER(u64)], i32),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    return L


def check(rc):
    if rc != 0:
        raise SvdagError(rc, lib().svdag_last_error().decode())


# --- HIP runtime (device buffers for the Python layer) ----------------------
HIP_H2D, HIP_D2H, HIP_D2D = 1, 2, 3


def hip():
    global _hip
    if _hip is None:
        H = None
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                H = ctypes.CDLL(name)
                break
            except OSError:
                H = None
        if H is None:
            raise RuntimeError("libamdhip64.so not found")
        _hip = bind_hip(H)
    return _hip


def bind_hip(H):
    """Declares the HIP runtime signatures used by the Python layer."""
    H.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    H.hipFree.argtypes = [ctypes.c_void_p]
    H.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    H.hipDeviceSynchronize.argtypes = []
    H.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
    H.hipGetErrorString.restype = ctypes.c_char_p
    H.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    H.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    H.hipEventSynchronize.argtypes = [ctypes.c_void_p]
    H.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
    H.hipEventDestroy.argtypes = [ctypes.c_void_p]
    H.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    H.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    H.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    H.hipMemcpyAsync.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p]
    return H


def hip_check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed: {hip().hipGetErrorString(rc).decode()}")


def gpu_count():
    n = ctypes.c_int(0)
    try:
        rc = hip().hipGetDeviceCount(ctypes.byref(n))
    except RuntimeError:
        return 0
    return n.value if rc == 0 else 0
