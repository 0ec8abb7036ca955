"""Seeded synthetic workloads of the benchmark configurations.

The reference builds its inputs with OpenGL from a glTF scene
(src/DAG_Compression_Fast.cpp:266-302, src/Voxelizer.cpp:254-369); these are
seeded stand-ins with the same input contract (unsorted uint64 Morton codes
with duplicates, per chunk for the chunked flows), generated on the host by the
library's generators (csrc/svdag_gen.cpp):

- spheres256: 64 sphere shells in a 256^3 grid (depth 8);
- triangles2048: 20k triangle surfaces in a 2048^3 grid (depth 11);
- noise1024: value noise thresholded at 0.5 in a 1024^3 grid (depth 10),
  emitted as 8 Morton-aligned 512^3 batches, each built on its own and merged
  (the reference's many-chunk flow); noise512 is one 512^3 cube of it with
  32-bit keys (depth 9);
- chunks512: 512 sub-volumes of 512^3 (chunk depth 9, global depth 12) with
  lognormal key counts, sphere-shell or triangle-surface content;
- served: 2000 such chunks as requests at distinct slots of a 4096-slot grid
  (global depth 13) with Poisson arrivals.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import gen_chunk, gen_chunk_plan, gen_value_noise

CHUNKS = dict(desc="512 sub-volumes of 512^3 (chunk depth 9 -> depth 12), lognormal keys per chunk "
                   "(median 2e5, sigma 1, 1e3..4e6), sphere-shell or triangle-surface content, 20-30% duplicates",
              seed=0x5EED0003, num=512, depth=9, levels_above=3)

SERVED = dict(desc="served 512^3 chunk requests (chunk depth 9) at distinct slots of a 16^3-slot grid (4096 slots, "
                   "depth 13), key counts and content drawn as in the 512-chunk workload, Poisson arrivals",
              seed=0x5EED0004, requests=2000, depth=9, slot_levels=4,
              rates=(50.0, 100.0, 200.0, 500.0, 1000.0, 2000.0), min_seconds=30.0)

NOISE = dict(desc="value noise 1024^3 (depth 10): lattice spacing 8, trilinear at voxel centres >= 0.5 in fixed "
                  "point, 8 Morton-aligned 512^3 batches (shuffled within a batch), each built at depth 9 and "
                  "merged to depth 10",
             seed=0x5EED0002, depth=10, spacing=8, batch_depth=9)

NOISE_U32 = dict(desc="value noise 512^3 (depth 9), 32-bit keys (27 significant bits), one shuffled batch",
                 seed=0x5EED0002, depth=9, spacing=8)


def chunk_workload(threads, seed=CHUNKS["seed"], num=CHUNKS["num"], depth=CHUNKS["depth"]):
    """The 512-chunk workload: one host array of chunk-local keys per chunk, in
    Morton order of the chunks."""
    n, _, _ = gen_chunk_plan(seed, num)
    with ThreadPoolExecutor(max(1, threads)) as ex:
        return list(ex.map(lambda c: gen_chunk(seed, c, depth, capacity=int(n[c])), range(num)))


def chunk_offsets(chunks):
    """Offsets of chunks laid back to back (num + 1 entries)."""
    offsets = np.zeros(len(chunks) + 1, dtype=np.uint64)
    offsets[1:] = np.cumsum([len(c) for c in chunks])
    return offsets


def chunk_union(chunks, chunk_depth):
    """Global keys of chunks in Morton order: chunk i's index above its
    chunk-local codes (src/Utils/Aabb.h:83-105)."""
    shift = np.uint64(3 * chunk_depth)
    return np.concatenate([(np.uint64(i) << shift) | np.asarray(c, dtype=np.uint64) for i, c in enumerate(chunks)])


def noise_batches(threads, spec=NOISE):
    """The 8 Morton-aligned batches of the 1024^3 value-noise grid, each as
    batch-local keys (< 2^(3 * batch_depth)), in Morton order."""
    sub = spec["batch_depth"]
    count = 1 << (3 * (spec["depth"] - sub))
    mask = np.uint64((1 << (3 * sub)) - 1)
    per = max(1, threads // count)

    def one(i):
        return gen_value_noise(spec["seed"], spec["depth"], spec["spacing"], sub, i, threads=per) & mask

    with ThreadPoolExecutor(count) as ex:
        return list(ex.map(one, range(count)))


def noise_u32(threads, spec=NOISE_U32):
    """One 512^3 value-noise cube as uint32 keys."""
    keys = gen_value_noise(spec["seed"], spec["depth"], spec["spacing"], spec["depth"], 0, threads=threads)
    return keys.astype(np.uint32)
