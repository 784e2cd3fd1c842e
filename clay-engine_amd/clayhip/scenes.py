"""Synthetic workloads of the fluid-to-field path, with their sizes.

Seeded jitter comes from the reference's own hash (math/gpu.ts:103-117), evaluated
here in exact u32 arithmetic, so every consumer of a scene (the GPU path, tests,
the CPU baseline) starts from bit-identical particle arrays.
"""
from __future__ import annotations

import math

import numpy as np

ORIGIN = (-12.0, -4.0, -12.0)
WORLD = 24.0


def hash_u32(x: np.ndarray) -> np.ndarray:
    """lowbias32 (math/gpu.ts:103-112) on uint32 arrays."""
    x = np.asarray(x, dtype=np.uint32).copy()
    x ^= x >> np.uint32(16)
    x *= np.uint32(0x7FEB352D)
    x ^= x >> np.uint32(15)
    x *= np.uint32(0x846CA68B)
    x ^= x >> np.uint32(16)
    return x


def hash1(seed: np.ndarray) -> np.ndarray:
    """hash1 -> [0, 1) as float32 (exact: 24 bits times 2^-24)."""
    return (hash_u32(seed) & np.uint32(0xFFFFFF)).astype(np.float32) * np.float32(1.0 / 16777216.0)


def hash3(seed: np.ndarray) -> np.ndarray:
    s = np.asarray(seed, dtype=np.uint32) * np.uint32(3)
    return np.stack([hash1(s), hash1(s + np.uint32(1)), hash1(s + np.uint32(2))], axis=-1)


def lattice(nx: int, ny: int, nz: int, spacing: float, lo, seed: int = 1,
            jitter: float = 0.1) -> np.ndarray:
    """A jittered simple-cubic block: particle n = ix + nx (iy + ny iz) sits at
    lo + (i + 0.5) spacing plus a uniform offset in +-jitter*spacing per axis."""
    iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    idx = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=-1).astype(np.float64)
    n = idx.shape[0]
    base = np.asarray(lo, dtype=np.float64) + (idx + 0.5) * spacing
    ids = np.arange(n, dtype=np.uint64)
    seeds = ((np.uint64(seed) * np.uint64(9781) + ids) & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    u = hash3(seeds).astype(np.float64)
    pos = base + (u * 2.0 - 1.0) * (jitter * spacing)
    return pos.astype(np.float32)


def floor_brush():
    """The roundBox floor: half-extents (11, 1, 11), corner 0.25, centred (0, -1.4, 0)."""
    return {"kind": "roundBox", "size": (11, 1, 11), "radius": 0.25, "pos": (0, -1.4, 0)}


def ridge_world_brushes():
    """Floor plus a box ridge (9.5, 1.4, 0.9) at (0, 0.4, 0) with a radius-2 sphere at
    (2.5, 0.8, 0) cut out, as the shape DSL flattens `union(floor, cut(box, sphere))`
    into a brush fold (lib/src/shape/sdf.ts:153-172)."""
    return [
        floor_brush(),
        {"kind": "box", "size": (9.5, 1.4, 0.9), "pos": (0, 0.4, 0)},
        {"kind": "sphere", "size": 2.0, "pos": (2.5, 0.8, 0), "op": "cut"},
    ]


def dam_break_options(capacity: int, bake_res: int, bake_mips: int) -> dict:
    """FluidSim options shared by the dam-break workloads: spacing 0.05, h 0.1,
    rest density 1000 (mass = 1000 spacing^3), g -9.81, dt 1/120 (2 substeps),
    collision radius 0.5 spacing, splat radius 0.05, band 4 over a 24-unit box."""
    return {
        "capacity": capacity,
        "spacing": 0.05,
        "smoothing": 0.1,
        "density": 1000.0,
        "substeps": 2,
        "gravity": (0.0, -9.81, 0.0),
        "surfaceRadius": 0.05,
        "bakeResolution": bake_res,
        "bakeMipLevels": bake_mips,
        "bakeBand": 4,
        "origin": ORIGIN,
        "worldSize": WORLD,
    }


def dam_break_small(seed: int = 1) -> dict:
    """16 384 particles (16 x 64 x 16 at 0.05 over [-0.4,0.4] x [0,3.2] x [-0.4,0.4]),
    analytic roundBox floor, 64^3 bake with 4 mips."""
    pos = lattice(16, 64, 16, 0.05, (-0.4, 0.0, -0.4), seed=seed)
    return {
        "name": "dam-break 16k, analytic floor, 64^3 x 4 mips",
        "options": dam_break_options(pos.shape[0], 64, 4),
        "collider": {"kind": "analytic", "brushes": [floor_brush()]},
        "pos": pos,
        "steps": 100,
    }


def dam_break_large(seed: int = 1) -> dict:
    """1 048 576 particles (128 x 64 x 128 at 0.05 over [-3.2,3.2] x [0,3.2] x [-3.2,3.2]),
    288^3 baked world collider (3 mips), 288^3 bake with 3 mips (the most 288 allows,
    volume.ts:156-160)."""
    pos = lattice(128, 64, 128, 0.05, (-3.2, 0.0, -3.2), seed=seed)
    return {
        "name": "dam-break 1M, 288^3 baked world, 288^3 x 3 mips",
        "options": dam_break_options(pos.shape[0], 288, 3),
        "collider": {
            "kind": "volume",
            "volume": {"resolution": 288, "mipLevels": 3, "worldSize": WORLD, "origin": ORIGIN, "band": 4},
            "brushes": ridge_world_brushes(),
        },
        "pos": pos,
        "steps": 100,
    }


def scaled_lattice_count(nx: int, ny: int, nz: int) -> int:
    return nx * ny * nz


def batched_counts(scenes: int = 256, lo: int = 1024, hi: int = 65536, seed: int = 7) -> np.ndarray:
    """Particle counts drawn log-uniformly in [lo, hi] from the reference hash."""
    u = hash1(np.arange(scenes, dtype=np.uint32) + np.uint32(seed * 7919)).astype(np.float64)
    return np.floor(np.exp(math.log(lo) + u * (math.log(hi) - math.log(lo)))).astype(np.int64)


BATCH_ORIGIN = (-3.2, -0.8, -3.2)
BATCH_WORLD = 6.4
BATCH_FOLLOW = 3.2


def batch_options(bake_res: int = 32, bake_mips: int = 3, follow: float = BATCH_FOLLOW) -> dict:
    """Options of every scene of the batched workload: the dam-break constants in a
    6.4-unit world over the same roundBox floor, each scene baking a 32^3 x 3-mip
    volume 3.2 units wide that follows its particles."""
    opts = dam_break_options(1, bake_res, bake_mips)
    opts.update({"origin": BATCH_ORIGIN, "worldSize": BATCH_WORLD, "followSize": follow,
                 "splatMode": "tiled"})
    return opts


def drop_block(count: int, drop: float, seed: int, spacing: float = 0.05) -> np.ndarray:
    """`count` particles of a jittered lattice block (nx = nz = ceil(cbrt(count)),
    filled layer by layer) centred on x = z = 0 with its base at height `drop`."""
    nx = int(math.ceil(round(count ** (1.0 / 3.0), 9)))
    while nx ** 3 < count:
        nx += 1
    ny = int(math.ceil(count / (nx * nx)))
    lo = (-nx * spacing / 2, drop, -nx * spacing / 2)
    return lattice(nx, ny, nx, spacing, lo, seed=seed)[:count]


def batched(scenes: int = 256, seed: int = 7) -> dict:
    """Independent scenes with particle counts log-uniform in [1024, 65536] and drop
    heights uniform in [0.5, 3.0], all from the reference hash; positions are the
    scenes' blocks concatenated scene-major."""
    counts = batched_counts(scenes, seed=seed)
    u = hash1(np.arange(scenes, dtype=np.uint32) + np.uint32(seed * 7919 + 104729)).astype(np.float64)
    drops = 0.5 + 2.5 * u
    blocks = [drop_block(int(counts[i]), float(drops[i]), seed=seed * 1000 + i) for i in range(scenes)]
    return {
        "name": f"batched {scenes} scenes, {int(counts.sum())} particles, following 32^3 x 3 mips each",
        "options": batch_options(),
        "collider": {"kind": "analytic", "brushes": [floor_brush()]},
        "counts": counts.astype(np.uint32),
        "drops": drops,
        "pos": np.concatenate(blocks, axis=0),
        "steps": 100,
    }


def batched_even(scenes: int = 64, count: int = 16384, seed: int = 7) -> dict:
    """`scenes` independent scenes of `count` particles each (64 x 16 384 = 1 048 576
    in total, the particle count of the single-scene 1M dam break), drop heights
    uniform in [0.5, 3.0] from the reference hash, with the batched workload's
    options."""
    counts = np.full(scenes, count, dtype=np.int64)
    u = hash1(np.arange(scenes, dtype=np.uint32) + np.uint32(seed * 7919 + 104729)).astype(np.float64)
    drops = 0.5 + 2.5 * u
    blocks = [drop_block(count, float(drops[i]), seed=seed * 1000 + i) for i in range(scenes)]
    return {
        "name": f"batched {scenes} scenes x {count} particles ({scenes * count} in total), following 32^3 x 3 mips each",
        "options": batch_options(),
        "collider": {"kind": "analytic", "brushes": [floor_brush()]},
        "counts": counts.astype(np.uint32),
        "drops": drops,
        "pos": np.concatenate(blocks, axis=0),
        "steps": 100,
    }
