"""Served mode: chunk-build requests with continuous admission.

Requests are chunks of voxel keys (host arrays, chunk-local Morton codes) that
arrive over time, each for a distinct sub-volume of a 8^k chunk grid (the
reference's per-chunk loop, src/DAG_Compression_Fast.cpp:266-330, fed by a
request stream).

- **Slots.** Each slot is a host thread with its own build context on its own
  HIP stream and its own device key buffer. A free slot takes the oldest
  arrived request, so the device work of admitted requests overlaps.
- **Admission.** A request is admitted only while the device bytes of the
  requests in flight (keys + build scratch, from svdag_scratch_bytes) stay
  within the HBM budget.
- **Results.** Each request's result is its chunk DAG, kept on the device; its
  latency runs from arrival to that DAG being complete (upload included). The
  global DAG is the merge of all chunk DAGs (DAGCompression::MergeDAGs), which
  equals the batched build of the same chunks.
"""
import collections
import ctypes
import threading
import time

import numpy as np

from . import _lib
from . import Context, Dag, DeviceBuffer

Request = collections.namedtuple("Request", "index chunk keys arrival")


class Result:
    __slots__ = ("index", "chunk", "dag", "n", "arrival", "start", "done", "slot")

    def __init__(self, req, dag, start, done, slot):
        self.index, self.chunk, self.n, self.arrival = req.index, req.chunk, len(req.keys), req.arrival
        self.dag, self.start, self.done, self.slot = dag, start, done, slot

    @property
    def latency(self):
        return self.done - self.arrival


class ChunkServer:
    """Continuous-admission server for chunk builds (see the module docstring)."""

    def __init__(self, chunk_depth, slots=4, budget_bytes=None, max_chunk_keys=1 << 22):
        self.chunk_depth = int(chunk_depth)
        self.slots = int(slots)
        self.max_chunk_keys = int(max_chunk_keys)
        per_slot = self.footprint(self.max_chunk_keys)
        self.budget = int(budget_bytes) if budget_bytes else per_slot * self.slots
        H = _lib.hip()
        self._streams = []
        self._ctx = []
        self._keys = []
        for _ in range(self.slots):
            s = ctypes.c_void_p()
            _lib.hip_check(H.hipStreamCreate(ctypes.byref(s)), "hipStreamCreate")
            self._streams.append(s)
            self._ctx.append(Context(s))
            self._keys.append(DeviceBuffer(8 * self.max_chunk_keys))

    def footprint(self, n):
        """Device bytes one request holds while in flight: its keys and build scratch."""
        return 8 * int(n) + int(_lib.lib().svdag_scratch_bytes(int(n), self.chunk_depth, 8))

    def run(self, requests):
        """Serves `requests` (sorted by arrival, seconds from the call) and returns
        the results in request order. Arrivals are replayed against the wall clock."""
        for r in requests:
            if len(r.keys) > self.max_chunk_keys:
                raise ValueError(f"chunk of {len(r.keys)} keys exceeds the slot capacity {self.max_chunk_keys}")
        lock = threading.Condition()
        pending = collections.deque(requests)
        state = {"in_flight": 0, "error": None}
        results = [None] * len(requests)
        t0 = time.perf_counter()
        L, H = _lib.lib(), _lib.hip()

        def take():
            with lock:
                while True:
                    if state["error"] or not pending:
                        return None
                    head = pending[0]
                    now = time.perf_counter() - t0
                    if head.arrival > now:
                        lock.wait(timeout=head.arrival - now)
                        continue
                    need = self.footprint(len(head.keys))
                    if state["in_flight"] and state["in_flight"] + need > self.budget:
                        lock.wait()
                        continue
                    pending.popleft()
                    state["in_flight"] += need
                    return head, need

        def slot_loop(slot):
            ctx, buf, stream = self._ctx[slot], self._keys[slot], self._streams[slot]
            while True:
                item = take()
                if item is None:
                    return
                req, need = item
                try:
                    start = time.perf_counter() - t0
                    keys = np.ascontiguousarray(req.keys, dtype=np.uint64)
                    if keys.size:
                        _lib.hip_check(H.hipMemcpyAsync(buf.ptr, keys.ctypes.data, keys.nbytes, _lib.HIP_H2D,
                                                        stream), "hipMemcpyAsync")
                    raw = _lib.SvdagDag()
                    _lib.check(L.svdag_create_dag(ctx.handle, buf.ptr, keys.size, self.chunk_depth,
                                                  ctypes.byref(raw)))
                    done = time.perf_counter() - t0
                    results[req.index] = Result(req, Dag(raw, ctx), start, done, slot)
                except Exception as e:  # noqa: BLE001 -- reported by run()
                    with lock:
                        state["error"] = e
                finally:
                    with lock:
                        state["in_flight"] -= need
                        lock.notify_all()

        threads = [threading.Thread(target=slot_loop, args=(i,), daemon=True) for i in range(self.slots)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        if state["error"] is not None:
            raise state["error"]
        return results

    def merge(self, results, num_chunks):
        """Global DAG of the served chunks (slots without a request are empty)."""
        empty = _lib.SvdagDag()
        empty.num_levels = self.chunk_depth - 2
        empty.location = _lib.SVDAG_DEVICE
        arr = (_lib.SvdagDag * num_chunks)()
        for i in range(num_chunks):
            arr[i] = empty
        for r in results:
            arr[r.chunk] = r.dag.raw
        out = _lib.SvdagDag()
        ctx = self._ctx[0]
        for s in self._streams:
            _lib.hip_check(_lib.hip().hipStreamSynchronize(s), "hipStreamSynchronize")
        _lib.check(_lib.lib().svdag_merge_dags(ctx.handle, arr, num_chunks, ctypes.byref(out)))
        return Dag(out, ctx)

    def close(self):
        for r in self._keys:
            r.free()
        for c in self._ctx:
            c.close()
        H = _lib.hip()
        for s in self._streams:
            H.hipStreamDestroy(s)
        self._streams, self._ctx, self._keys = [], [], []


def poisson_arrivals(rate, count, seed):
    """Arrival times (seconds) of a Poisson process of `rate` requests/s."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=count))


def latency_stats(results):
    lat = np.array([r.latency for r in results])
    return {"p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "max_ms": float(lat.max() * 1e3)}
