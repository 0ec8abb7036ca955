"""The product sources under AddressSanitizer and UndefinedBehaviorSanitizer.

tests/emu builds the library's .hip sources for the CPU kernel emulator with
-fsanitize=address,undefined (make sanitize).  A child process with the ASan
runtime preloaded runs builds through the C ABI over the paths the reference
has device bounds checks for (its ENABLE_CHECKS kernels, DAGCompression.cu:
278-334): plain builds at several depths and sizes on both key-sort routes,
32-bit keys, batched builds (per-chunk and global sort, up to 4096 chunks),
merges, the flat layout, copies to the
host and back, the resident DAG, a captured batched build, the no-allocation
entry point, the key sort, point queries and the voxelizer.  Any sanitizer
report fails the test."""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))

SCRIPT = r'''
import sys
sys.path[:0] = [{tests!r}, {emu!r}, {oracle!r}, {root!r}]
import numpy as np
import emu_ctypes
import oracle_ctypes as oracle
from dagcheck import assert_dag_equal
import test_gpu_parity as T
m = emu_ctypes.use_emulator()
rng = np.random.default_rng(42)
for depth, n in ((3, 1), (4, 300), (7, 5000), (10, 20000), (21, 2000)):
    keys = rng.integers(0, 1 << (3 * depth), size=n, dtype=np.uint64)
    T.check_case(m, oracle, np.concatenate([keys, keys[: n // 3]]), depth, "sanitized")
T.check_case(m, oracle, T.clustered_keys(rng, 9, 20000), 9, "clustered")
T.check_case(m, oracle, T.solid_keys(rng, 7, 500), 7, "solid")
u32 = m.create_dag(np.arange(3000, dtype=np.uint32), 8)
assert_dag_equal(u32.to_numpy(), oracle.create_dag(np.arange(3000, dtype=np.uint64), 8), "u32")
import test_top_digit_first as TD
for depth, n in ((9, 3000), (11, 4000), (12, 3000)):  # one and two segmented passes
    keys = rng.integers(0, 1 << (3 * depth), size=n, dtype=np.uint64)
    TD.built(m, oracle, keys, depth, TD.ALWAYS)
T.test_merge_dags_matches_oracle_and_global(m, oracle)
T.test_batched_many_sparse_chunks(m, oracle)
T.test_batched_chunk_depths(m, oracle, 8)
T.test_batched_chunk_depths(m, oracle, 10)
d = m.create_dag(T.clustered_keys(rng, 8, 5000), 8)
assert np.array_equal(m.create_final_dag(d.to_host().to_device()), m.create_final_dag(d))
import test_resident, test_graph, test_zz_no_alloc, test_sort_keys
test_resident.run_rounds(m, oracle)
test_resident.run_gather(m, oracle)
test_graph.run_graph(m, oracle, 6, 64, 2000, replays=1)
test_zz_no_alloc.run_cases(m)
test_zz_no_alloc.run_errors(m)
for n, bits in ((4097, 33), (5000, 64)):
    test_sort_keys.check_sort(m, rng, n, bits)
words = m.create_final_dag(d)
codes = rng.integers(0, 8 ** 8, size=2000, dtype=np.uint64)
m.query_voxels(words, 8, codes)
tris = m.gen_triangles(5, 6, 40, 1.5, 0.5)
m.voxelize_triangles(tris, 6)
print("SANITIZED-OK", flush=True)
'''


def test_product_sources_under_sanitizers():
    emu = os.path.join(HERE, "emu")
    subprocess.run(["make", "-s", "-C", emu, "sanitize"], check=True)
    lib = os.path.join(emu, "_build", "libsvdag_emu_san.so")
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    if not os.path.isabs(asan) or not os.path.exists(asan):
        pytest.skip("no ASan runtime for the host compiler")
    root = os.path.dirname(HERE)
    script = SCRIPT.format(tests=HERE, emu=emu, oracle=os.path.join(root, "oracle"), root=root)
    env = dict(os.environ)
    env["SVDAG_EMU_LIB"] = lib
    env["LD_PRELOAD"] = asan + (":" + env["LD_PRELOAD"] if env.get("LD_PRELOAD") else "")
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=0:exitcode=23"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    res = subprocess.run([sys.executable, "-c", script], env=env, capture_output=True, text=True, timeout=1200)
    report = res.stdout[-3000:] + res.stderr[-6000:]
    assert res.returncode == 0 and "SANITIZED-OK" in res.stdout, report
    assert "AddressSanitizer" not in res.stderr and "runtime error" not in res.stderr, report
