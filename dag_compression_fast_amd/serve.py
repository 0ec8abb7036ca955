"""Served mode: chunk-compression requests with continuous batching into a
global DAG resident on the device.

Each request is one chunk (a 2^chunk_depth sub-volume at its own slot of an
8^slot_levels grid of slots) given as unsorted chunk-local Morton keys, either
already in device memory (a DeviceBuffer, as the reference's voxelizer leaves
its fragments in a GPU buffer, src/DAG_Compression_Fast.cpp:283-302) or as a
host array uploaded on admission.  The server replays the arrivals against
the wall clock:

- **Continuous batching.**  Whenever the device is free, every request that
  has arrived and fits the HBM budget joins the next round; requests that
  arrive while a round is on the device wait for the following one.  A round
  is one batched build over the slots of its requests (their work runs in the
  same kernels), then one insertion into the resident DAG.
- **Admission by HBM footprint.**  A round's footprint is its keys, the build's
  scratch for them (svdag_scratch_bytes) and its result; requests are admitted
  in arrival order while the round stays within the budget.  The resident DAG's
  tables are sized once (max_nodes distinct nodes per level); a round that
  could overflow them is refused before any device work.
- **Resident state.**  The global DAG and its per-level dedup tables stay on
  the device between rounds (svdag_resident_*); each round adds exactly the
  nodes the union gains.  ``dag()`` returns the global DAG, equal to the
  reference's MergeDAGs over every served chunk (src/DAGCompression.cu:1447-1548).
- **Results.**  A request's result is its chunk's root hash in the global DAG
  (the content identity of its chunk DAG: the root hash of the chunk built
  alone), and its latency runs from arrival to the end of its round's
  insertion (upload included for host keys).
"""
import collections
import time

import numpy as np

from . import _lib
from . import Context, DeviceBuffer, ResidentDag, create_dag_batched_gather

Request = collections.namedtuple("Request", "index slot keys arrival")


class Result:
    __slots__ = ("index", "slot", "root_hash", "n", "arrival", "start", "done", "round")

    def __init__(self, req, n, root_hash, start, done, rnd):
        self.index, self.slot, self.n, self.arrival = req.index, req.slot, n, req.arrival
        self.root_hash, self.start, self.done, self.round = root_hash, start, done, rnd

    @property
    def latency(self):
        return self.done - self.arrival


def _count(keys):
    return int(keys.count) if isinstance(keys, DeviceBuffer) else len(keys)


class ChunkServer:
    """Continuous-batching server over a resident global DAG (module docstring)."""

    def __init__(self, chunk_depth, slot_levels, max_nodes, budget_bytes=None, ctx=None, clock=time.perf_counter):
        self.chunk_depth = int(chunk_depth)
        self.slot_levels = int(slot_levels)
        self.num_slots = 8 ** self.slot_levels
        self.depth = self.chunk_depth + self.slot_levels
        self.ctx = ctx or Context()
        self._own_ctx = ctx is None
        self.clock = clock
        self.resident = ResidentDag(self.chunk_depth, self.slot_levels, max_nodes, ctx=self.ctx)
        self.budget = int(budget_bytes) if budget_bytes else self.footprint(1 << 27)
        self.rounds = 0

    def footprint(self, n):
        """Device bytes a round of n keys holds while it runs: its keys (and
        their gathered copy), the build's scratch and its result."""
        L = _lib.lib()
        n = int(n)
        return 16 * n + int(L.svdag_scratch_bytes(n, self.depth, 8)) + int(L.svdag_result_bytes(n, self.depth))

    def _round(self, batch):
        """One round: the batched build of the requests' chunks at their slots,
        then its insertion into the resident DAG.  Returns the root hashes."""
        bufs = {}
        staged = []
        try:
            for r in batch:
                if isinstance(r.keys, DeviceBuffer):
                    bufs[r.slot] = r.keys
                elif len(r.keys):
                    b = DeviceBuffer.from_numpy(np.ascontiguousarray(r.keys, dtype=np.uint64))
                    staged.append(b)
                    bufs[r.slot] = b
            rd = create_dag_batched_gather(bufs, self.chunk_depth, ctx=self.ctx, num_chunks=self.num_slots)
        finally:
            for b in staged:
                b.free()
        try:
            return self.resident.insert(rd)
        finally:
            rd.free()

    def run(self, requests):
        """Serves `requests` (arrival times in seconds from the call) and returns
        the results in request order."""
        seen = set()
        for r in requests:
            if not 0 <= r.slot < self.num_slots or r.slot in seen:
                raise ValueError(f"request {r.index}: slot {r.slot} out of range or used twice")
            seen.add(r.slot)
            if self.footprint(_count(r.keys)) > self.budget:
                raise ValueError(f"request {r.index} ({_count(r.keys)} keys) exceeds the HBM budget")
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        results = [None] * len(requests)
        t0 = self.clock()
        while pending:
            now = self.clock() - t0
            if pending[0].arrival > now:
                # Nothing has arrived: wait for the next arrival.
                time.sleep(min(pending[0].arrival - now, 0.0005))
                continue
            batch, keys = [], 0
            while pending and pending[0].arrival <= now:
                n = _count(pending[0].keys)
                if batch and self.footprint(keys + n) > self.budget:
                    break
                batch.append(pending.popleft())
                keys += n
            start = self.clock() - t0
            hashes = self._round(batch)
            done = self.clock() - t0
            for r in batch:
                results[r.index] = Result(r, _count(r.keys), int(hashes[r.slot]), start, done, self.rounds)
            self.rounds += 1
        return results

    def dag(self):
        """The global DAG of every chunk served so far (levels in hash order)."""
        return self.resident.dag()

    def close(self):
        self.resident.close()
        if self._own_ctx:
            self.ctx.close()


def poisson_arrivals(rate, count, seed):
    """Arrival times (seconds) of a Poisson process of `rate` requests/s."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=count))


def latency_stats(results):
    lat = np.array([r.latency for r in results])
    return {"p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "max_ms": float(lat.max() * 1e3)}
