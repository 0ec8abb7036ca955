"""Committed golden vectors (tests/golden/svdag_golden.npz, written and
cross-checked by tests/golden/make_golden.py): the oracle, the CPU kernel
emulator and the GPU library must reproduce every case bit for bit."""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "svdag_golden.npz")

def load():
    with np.load(GOLDEN) as z:
        data = {k: z[k] for k in z.files}
    names = sorted({k.split("/")[0] for k in data})
    return [(n, data[f"{n}/keys"], int(data[f"{n}/depth"][0]), data[f"{n}/level_counts"],
             int(data[f"{n}/num_leaves"][0]), data[f"{n}/words"]) for n in names]


CASES = load()


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_oracle_reproduces_golden(oracle, case):
    name, keys, depth, counts, leaves, words = case
    d = oracle.create_dag(keys, depth)
    assert [len(lv["hashes"]) for lv in d["levels"]] == counts.tolist()
    assert len(d["leaves"]) == leaves
    assert np.array_equal(oracle.create_final_dag(d), words)


@pytest.fixture(scope="module")
def emu():
    import emu_ctypes
    svd = emu_ctypes.use_emulator()
    yield svd
    from dag_compression_fast_amd import _lib
    _lib._lib = None
    _lib._hip = None
    svd._default = None


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_emulator_reproduces_golden(emu, case):
    name, keys, depth, counts, leaves, words = case
    dag = emu.create_dag(keys, depth)
    assert dag.level_counts == counts.tolist()
    assert dag.num_leaves == leaves
    assert np.array_equal(emu.create_final_dag(dag), words)


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_gpu_reproduces_golden(gpu, case):
    name, keys, depth, counts, leaves, words = case
    print(f"[golden] {name}: depth={depth} n={len(keys)}", flush=True)
    dag = gpu.create_dag(keys, depth)
    assert dag.level_counts == counts.tolist()
    assert dag.num_leaves == leaves
    assert np.array_equal(gpu.create_final_dag(dag), words)
