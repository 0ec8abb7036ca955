import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests"), os.path.join(ROOT, "tests", "emu")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU; runs the HIP library")
    config.addinivalue_line("markers", "slow: full-size workloads")


@pytest.fixture(scope="session")
def oracle():
    import oracle_ctypes
    oracle_ctypes.lib()
    return oracle_ctypes


@pytest.fixture(scope="session")
def gpu():
    """The product package with a visible GPU; fails (never skips) without one."""
    import dag_compression_fast_amd as svd
    from dag_compression_fast_amd import _lib
    if _lib.gpu_count() < 1:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")
    # The library fails a call whose device work does not finish in time and
    # names the stage; keep that well inside the per-test timeout.
    svd.default_context().set_option(_lib.OPT_WAIT_SECONDS, 60)
    return svd
