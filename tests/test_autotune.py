"""Autotuner: picks the fastest setting, leaves it applied, refuses a setting
that changes the DAG, and keeps its choice in the JSON cache.  The CPU tests
drive it with a stand-in context and a stand-in clock; the GPU test tunes the
batched build of a small uneven-chunk workload and checks that the chosen
setting gives the same DAG as the defaults."""
import numpy as np
import pytest

from dag_compression_fast_amd import _lib, autotune


class FakeCtx:
    def __init__(self):
        self.options = {}

    def set_option(self, option, value):
        self.options[option] = value


def setting_of(ctx):
    names = {v: k for k, v in autotune._OPTION_IDS.items()}
    return {names[o]: v for o, v in ctx.options.items()}


def fake_cost(setting):
    # Fastest at radix 8, tile 4096, top digit first, load shift 2 (made-up costs).
    return (abs(setting["radix_bits"] - 8) + (setting["sort_tile"] == 8192) * 2
            + (setting["table_load_shift"] == 1) * 0.5 + (setting["top_digit_first"] == 0) * 0.25 + 1.0)


def make_run(ctx, words_for=lambda s: np.arange(10, dtype=np.uint32)):
    now = [0.0]

    def clock():
        return now[0]

    def build(c):
        now[0] += fake_cost(setting_of(c)) / 1e3

    def fetch():
        return words_for(setting_of(ctx))

    return build, fetch, clock


def test_picks_fastest_and_applies_it():
    ctx = FakeCtx()
    build, fetch, clock = make_run(ctx)
    res = autotune.autotune(ctx, build, fetch, reps=2, clock=clock)
    assert res["best"] == {"radix_bits": 8, "sort_tile": 4096, "table_load_shift": 2, "top_digit_first": 1}
    assert setting_of(ctx) == res["best"]
    assert len(res["table"]) == 3 * 2 * 2 * 2
    assert res["best_ms"] == pytest.approx(1.0)


def test_setting_that_changes_the_dag_is_refused():
    ctx = FakeCtx()

    def words_for(s):
        w = np.arange(10, dtype=np.uint32)
        if s["radix_bits"] == 8 and s["sort_tile"] == 8192:
            w[3] += 1
        return w

    build, fetch, clock = make_run(ctx, words_for)
    with pytest.raises(autotune.TuneMismatch):
        autotune.autotune(ctx, build, fetch, reps=1, clock=clock)


def test_cache_round_trip(tmp_path):
    ctx = FakeCtx()
    build, fetch, clock = make_run(ctx)
    res = autotune.autotune(ctx, build, fetch, reps=1, clock=clock)
    path = str(tmp_path / "tune.json")
    assert autotune.load_cached("gfx950|chunks", path) is None
    autotune.store("gfx950|chunks", res, path)
    assert autotune.load_cached("gfx950|chunks", path) == res["best"]
    assert autotune.load_cached("other", path) is None


def test_option_ids_match_the_header():
    text = open(_lib.HEADER).read()
    for name, opt in (("RADIX_BITS", _lib.OPT_RADIX_BITS), ("SORT_TILE", _lib.OPT_SORT_TILE),
                      ("TABLE_LOAD_SHIFT", _lib.OPT_TABLE_LOAD_SHIFT)):
        assert f"SVDAG_OPT_{name} = {opt}," in text
    assert f"SVDAG_OPT_SEG_FIRST_BITS = {_lib.OPT_SEG_FIRST_BITS}," in text
    assert f"SVDAG_OPT_TOP_DIGIT_FIRST = {_lib.OPT_TOP_DIGIT_FIRST} " in text


def test_space_follows_the_build_kind():
    # Chunk depth 9 leaves 12 chunk-local bits to sort: the per-chunk sort.
    assert autotune.space_for(9) is autotune.SEG_SPACE
    assert autotune.space_for(6) is autotune.SEG_SPACE
    assert autotune.space_for(10) is autotune.SPACE
    assert autotune.space_for(None) is autotune.SPACE
    assert autotune.default_setting(autotune.SEG_SPACE, 9) == {"seg_first_bits": 6, "table_load_shift": 1}


def test_default_time_is_reported():
    ctx = FakeCtx()
    build, fetch, clock = make_run(ctx)
    res = autotune.autotune(ctx, build, fetch, reps=1, clock=clock)
    assert res["default_ms"] == pytest.approx(fake_cost(autotune.default_setting(autotune.SPACE)))


@pytest.mark.gpu
def test_autotune_batched_build_on_gpu(gpu):
    rng = np.random.default_rng(21)
    depth, k = 6, 1
    chunks = [rng.integers(0, 8 ** depth, size=int(rng.integers(100, 20000)), dtype=np.uint64)
              for _ in range(8 ** k)]
    offsets = np.zeros(len(chunks) + 1, dtype=np.uint64)
    offsets[1:] = np.cumsum([len(c) for c in chunks])
    keys = np.concatenate(chunks)
    ctx = gpu.Context()
    try:
        want = gpu.create_final_dag(gpu.create_dag_batched(keys, offsets, depth, ctx=ctx))
        state = {}

        def build(c):
            state["words"] = gpu.create_final_dag(gpu.create_dag_batched(keys, offsets, depth, ctx=c))

        res = autotune.autotune(ctx, build, lambda: state["words"], reps=1,
                                space={"radix_bits": (6, 9), "sort_tile": (4096,), "table_load_shift": (1, 2)})
        assert len(res["table"]) == 4
        assert np.array_equal(gpu.create_final_dag(gpu.create_dag_batched(keys, offsets, depth, ctx=ctx)), want)
        # Chunk depth 6 sorts inside each chunk (3 bits): its digit split and the load factor.
        res = autotune.autotune(ctx, build, lambda: state["words"], reps=1, space=autotune.space_for(depth),
                                chunk_depth=depth)
        assert len(res["table"]) == 5 * 3
        assert np.array_equal(gpu.create_final_dag(gpu.create_dag_batched(keys, offsets, depth, ctx=ctx)), want)
    finally:
        ctx.close()
