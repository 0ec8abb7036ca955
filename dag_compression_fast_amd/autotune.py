"""Autotuner for the build's speed options.

The key-sort digit width (``SVDAG_OPT_RADIX_BITS``), the key-sort tile
(``SVDAG_OPT_SORT_TILE``), the key sort's route (``SVDAG_OPT_TOP_DIGIT_FIRST``),
the digit split of the batched build's per-chunk sort
(``SVDAG_OPT_SEG_FIRST_BITS``) and the dedup table load factor
(``SVDAG_OPT_TABLE_LOAD_SHIFT``) change how fast a build runs, never what it
produces.  ``autotune`` times a caller-supplied build under every combination
of a search space on one context, checks that every combination gives the same
flat layout as the first, leaves the fastest combination set on the context and
returns the whole table.  Results can be kept in a JSON cache keyed by device
and workload, so a later run starts from the stored choice.

Which options matter depends on the build: a plain build (and a batched build
whose chunk-local bits above the leaf groups exceed 12, chunk depth >= 10)
runs the global key sort, whose digit width, tile and route are tuned; a batched build
of chunk depth <= 9 sorts inside each chunk, where only the digit split of that
sort and the table load factor have an effect (``space_for``).

The reference has no counterpart: its sort and table parameters are
compile-time constants (src/Core.h:3-18, src/DAGCompression.cu:336-394).
"""
import itertools
import json
import os
import time

import numpy as np

from . import _lib

SPACE = {
    "radix_bits": (7, 8, 9),
    "sort_tile": (4096, 8192),
    "top_digit_first": (0, 1),
    "table_load_shift": (1, 2),
}

# Batched builds sorted inside each chunk: the first pass's digit width of that
# sort, and the table load factor.
SEG_SPACE = {
    "seg_first_bits": (4, 5, 6, 7, 8),
    "table_load_shift": (1, 2, 3),
}

_OPTION_IDS = {
    "radix_bits": _lib.OPT_RADIX_BITS,
    "sort_tile": _lib.OPT_SORT_TILE,
    "table_load_shift": _lib.OPT_TABLE_LOAD_SHIFT,
    "seg_first_bits": _lib.OPT_SEG_FIRST_BITS,
    "top_digit_first": _lib.OPT_TOP_DIGIT_FIRST,
}

DEFAULTS = {"radix_bits": 9, "sort_tile": 4096, "table_load_shift": 1, "seg_first_bits": 0, "top_digit_first": 1}

# The key sort skips a leaf's 6 voxel bits and its 9-bit place in its block of
# 512 leaves; the batched build sorts inside each chunk when at most this many
# chunk-local bits remain.
_SORT_SKIP_BITS = 15
_SEG_MAX_BITS = 12


def segmented(chunk_depth):
    """Whether a batched build of this chunk depth sorts inside each chunk."""
    bits = 3 * int(chunk_depth) - _SORT_SKIP_BITS
    return 0 < bits <= _SEG_MAX_BITS


def space_for(chunk_depth=None):
    """The search space that matters for a batched build of `chunk_depth`
    (None: a plain build)."""
    return SEG_SPACE if chunk_depth is not None and segmented(chunk_depth) else SPACE


def default_setting(space, chunk_depth=None):
    """The library's default setting expressed in `space`'s options (the
    segmented sort's default split is half its bits)."""
    out = {}
    for name in space:
        v = DEFAULTS[name]
        if name == "seg_first_bits" and v == 0 and chunk_depth is not None:
            v = (3 * int(chunk_depth) - _SORT_SKIP_BITS + 1) // 2
        out[name] = v
    return out


class TuneMismatch(RuntimeError):
    """A candidate setting produced a different DAG than the first one."""


def apply(ctx, setting):
    """Set every option of `setting` ({name: value}) on the context."""
    for name, value in setting.items():
        ctx.set_option(_OPTION_IDS[name], int(value))


def candidates(space=None):
    space = space or SPACE
    names = sorted(space)
    for values in itertools.product(*(space[n] for n in names)):
        yield dict(zip(names, values))


def autotune(ctx, build, fetch, reps=3, space=None, clock=time.perf_counter, chunk_depth=None):
    """Time `build(ctx)` (one complete build on the context, synchronised, its
    result kept by the callee) under every candidate setting; `fetch()` returns
    the last build's flat layout as a numpy array for the equality check.
    Returns {"best": setting, "best_ms": ms, "table": [(setting, ms), ...],
    "default_ms": ms of the library's default setting when it is a candidate}
    and leaves the best setting applied to ctx."""
    table = []
    reference = None
    for setting in candidates(space):
        apply(ctx, setting)
        build(ctx)  # warm-up (also sizes the context's scratch for this setting)
        words = np.asarray(fetch())
        if reference is None:
            reference = words.copy()
        elif not np.array_equal(words, reference):
            raise TuneMismatch(f"setting {setting} changed the DAG")
        best = None
        for _ in range(reps):
            t0 = clock()
            build(ctx)
            dt = clock() - t0
            best = dt if best is None else min(best, dt)
        table.append((setting, best * 1e3))
    best_setting, best_ms = min(table, key=lambda row: row[1])
    apply(ctx, best_setting)
    want = default_setting(space or SPACE, chunk_depth)
    default_ms = next((ms for st, ms in table if st == want), None)
    return {"best": best_setting, "best_ms": best_ms, "table": table, "default_ms": default_ms}


def cache_path():
    return os.environ.get("SVDAG_TUNE_CACHE") or os.path.join(os.path.expanduser("~"), ".cache",
                                                              "svdag_autotune.json")


def load_cached(key, path=None):
    """The stored setting for `key`, or None."""
    path = path or cache_path()
    try:
        with open(path) as f:
            entry = json.load(f).get(key)
    except (OSError, ValueError):
        return None
    if not entry or not entry.get("best") or not set(entry["best"]) <= set(_OPTION_IDS):
        return None
    return entry["best"]


def store(key, result, path=None):
    """Keep `result` (from autotune) under `key` in the JSON cache."""
    path = path or cache_path()
    try:
        with open(path) as f:
            data = json.load(f)
    except (OSError, ValueError):
        data = {}
    data[key] = {"best": result["best"], "best_ms": result["best_ms"],
                 "table": [[s, ms] for s, ms in result["table"]]}
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)
