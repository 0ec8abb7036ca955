"""Autotuner for the build's speed options.

The key-sort digit width (``SVDAG_OPT_RADIX_BITS``), the key-sort tile
(``SVDAG_OPT_SORT_TILE``) and the dedup table load factor
(``SVDAG_OPT_TABLE_LOAD_SHIFT``) change how fast a build runs, never what it
produces.  ``autotune`` times a caller-supplied build under every combination
of a search space on one context, checks that every combination gives the same
flat layout as the first, leaves the fastest combination set on the context and
returns the whole table.  Results can be kept in a JSON cache keyed by device
and workload, so a later run starts from the stored choice.

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
    "table_load_shift": (1, 2),
}

_OPTION_IDS = {
    "radix_bits": _lib.OPT_RADIX_BITS,
    "sort_tile": _lib.OPT_SORT_TILE,
    "table_load_shift": _lib.OPT_TABLE_LOAD_SHIFT,
}

DEFAULTS = {"radix_bits": 9, "sort_tile": 4096, "table_load_shift": 1}


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


def autotune(ctx, build, fetch, reps=3, space=None, clock=time.perf_counter):
    """Time `build(ctx)` (one complete build on the context, synchronised, its
    result kept by the callee) under every candidate setting; `fetch()` returns
    the last build's flat layout as a numpy array for the equality check.
    Returns {"best": setting, "best_ms": ms, "table": [(setting, ms), ...]}
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
    return {"best": best_setting, "best_ms": best_ms, "table": table}


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
    if not entry or set(entry.get("best", {})) != set(SPACE):
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
