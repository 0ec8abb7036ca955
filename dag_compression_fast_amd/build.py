"""Builds the gfx950 shared library ``_build/libsvdag.so`` in-tree.

Every HIP translation unit is compiled with hipcc for gfx950 only; the host
generator is plain C++.  Packed-FP32 instruction selection is disabled for the
device code (``-Xclang -target-feature -Xclang -packed-fp32-ops``).
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(OUT, "libsvdag.so")
INCLUDE = os.path.join(ROOT, "include")

HIP_SOURCES = ["svdag_sort.hip", "svdag_build.hip", "svdag_api.hip", "svdag_voxelize.hip", "svdag_trace.hip"]
CXX_SOURCES = ["svdag_gen.cpp"]
HEADERS = ["svdag_device.hpp", "svdag_internal.hpp"]

ARCH = "gfx950"


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def hip_flags():
    return [
        f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", INCLUDE, "-I", CSRC,
    ]


def _run(cmd):
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        raise RuntimeError("command failed: " + " ".join(cmd) + "\n" + res.stdout)
    return res.stdout


def build(verbose=False, jobs=None):
    """Compile every source (always) and link libsvdag.so; returns its path."""
    os.makedirs(OUT, exist_ok=True)
    hipcc = _hipcc()
    cmds = []
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(OUT, src.replace(".hip", ".o"))
        cmds.append([hipcc, *hip_flags(), "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in CXX_SOURCES:
        obj = os.path.join(OUT, src.replace(".cpp", ".o"))
        cmds.append(["g++", "-O3", "-march=x86-64-v3", "-ffp-contract=off", "-std=c++17", "-fPIC", "-Wall", "-pthread",
                     "-I", INCLUDE, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs or min(8, len(cmds))) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
