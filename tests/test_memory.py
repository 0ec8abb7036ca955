"""Device memory limits, forced from the host with SVDAG_OPT_MEMORY_BUDGET.

The reference's allocator frees its cached blocks and retries once when the
device is full, and aborts if the retry fails too (src/Memory.cpp:56-70,
src/DAGCompression.cu:85-89, 178-189).  Here:
- a build whose scratch exceeds the budget fails with SVDAG_ERR_MEMORY and
  the context stays usable: after the budget is raised the same build
  succeeds, bit-exact with the oracle;
- freed result blocks are cached for reuse; an allocation that would pass the
  budget releases them and is retried (the fallback), and succeeds when the
  live memory fits;
- when the live memory does not fit even then, the call fails with
  SVDAG_ERR_MEMORY and leaves the context usable."""
import numpy as np
import pytest

from dagcheck import assert_dag_equal

ERR_MEMORY = 3


def run_budget(m, oracle, n_big, n_small):
    import test_gpu_parity as T
    from dag_compression_fast_amd import _lib
    L = _lib.lib()
    rng = np.random.default_rng(17)
    big = T.clustered_keys(rng, 9, n_big)
    small = T.clustered_keys(rng, 9, n_small)
    need_big = int(L.svdag_scratch_bytes(len(big), 9, 8))
    ctx = m.Context()
    try:
        # Scratch past the budget: an error, and the context stays usable.
        ctx.set_option(_lib.OPT_MEMORY_BUDGET, need_big // 2)
        with pytest.raises(m.SvdagError) as e:
            m.create_dag(big, 9, ctx=ctx)
        assert e.value.code == ERR_MEMORY
        ctx.set_option(_lib.OPT_MEMORY_BUDGET, 0)
        d_big = m.create_dag(big, 9, ctx=ctx)
        assert_dag_equal(d_big.to_numpy(), oracle.create_dag(big, 9), "after raising the budget")
        use = ctx.memory_usage()
        assert use["results"] > 0 and use["cached"] == 0
        d_big.free()
        use = ctx.memory_usage()
        assert use["results"] == 0 and use["cached"] > 0
        # The small build needs a smaller block than the cached one (no reuse):
        # within a budget that leaves no room for both, the cached block is
        # released and the allocation retried.
        small_block = int(L.svdag_result_bytes(len(small), 9))
        assert 2 * small_block < use["cached"]
        budget = use["scratch"] + small_block + (1 << 20)
        assert use["scratch"] + use["cached"] + small_block > budget
        ctx.set_option(_lib.OPT_MEMORY_BUDGET, budget)
        d_small = m.create_dag(small, 9, ctx=ctx)
        assert_dag_equal(d_small.to_numpy(), oracle.create_dag(small, 9), "after releasing the cache")
        use = ctx.memory_usage()
        assert use["cached"] == 0 and use["results"] > 0
        # Live memory past the budget even without the cache: an error.
        with pytest.raises(m.SvdagError) as e:
            m.create_dag(small[: len(small) // 2], 9, ctx=ctx)
        assert e.value.code == ERR_MEMORY
        d_small.free()
        d_half = m.create_dag(small[: len(small) // 2], 9, ctx=ctx)
        assert_dag_equal(d_half.to_numpy(), oracle.create_dag(small[: len(small) // 2], 9), "after freeing")
        d_half.free()
        ctx.free_all()
        use = ctx.memory_usage()
        assert use == {"scratch": 0, "results": 0, "cached": 0}
    finally:
        ctx.close()


@pytest.fixture(scope="module")
def emu():
    import emu_ctypes
    svd = emu_ctypes.use_emulator()
    yield svd
    from dag_compression_fast_amd import _lib
    _lib._lib = None
    _lib._hip = None
    svd._default = None


def test_memory_budget_on_the_cpu_emulator(emu, oracle):
    run_budget(emu, oracle, 200000, 20000)


@pytest.mark.gpu
def test_memory_budget_on_gpu(gpu, oracle):
    run_budget(gpu, oracle, 2000000, 100000)
