"""Runs the product Python API on the CPU kernel emulator (tests only).

``use_emulator()`` builds tests/emu/_build/libsvdag_emu.so (the product .hip
sources compiled by g++ against the mock HIP header) and binds the package's
ctypes layer to it, so host logic and kernel logic can be exercised without a
GPU.  Nothing in the product imports this module.
"""
import ctypes
import fcntl
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
EMU_LIB = os.environ.get('SVDAG_EMU_LIB', os.path.join(HERE, '_build', 'libsvdag_emu.so'))


def build():
    # Test workers (pytest -n) may call this at once: one make at a time.
    os.makedirs(os.path.join(HERE, '_build'), exist_ok=True)
    with open(os.path.join(HERE, '_build', '.lock'), 'w') as lock:
        fcntl.flock(lock, fcntl.LOCK_EX)
        subprocess.run(["make", "-s", "-C", HERE], check=True)


def use_emulator():
    import dag_compression_fast_amd as svd
    from dag_compression_fast_amd import _lib
    build()
    L = ctypes.CDLL(EMU_LIB)
    _lib._lib = _lib.bind_library(L)
    _lib._hip = _lib.bind_hip(L)
    svd._default = None
    return svd
