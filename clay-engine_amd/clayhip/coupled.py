"""A coupled frame over batched scenes: fluid, a clay body and an eroding world.

Per scene, every frame (SoftBody.simulate, softbody.ts:405-429, then the fluid and
the erosion it drives):
  1. the body steps in a ClaySolver (pbd.ts:400-415) that holds every scene's body;
     body s collides with its scene's world instance,
  2. the body centroids are reduced in FP64 on the device (ClaySolver.match's
     centroid, pbd.ts:240-266) and the body's 32^3 volume re-centres on them
     (surface.setCenter(position), softbody.ts:424-428) and bakes,
  3. the fluid steps against union(world instance, body volume) (field.ts:246-315),
  4. its contacts become cut edits of the scene's world instance, whose touched
     tiles and their coarser mips are rebuilt (SdfEditor.flush, modify.ts:214-238).
All of it runs on one stream with no host readback; scenes share every launch.
The solver takes SoftBody's settings (softbody.ts:228-233, 285-297): stiffness
0.55, plasticity 0.35 per second as a per-substep rate, gravity -18, friction
0.45, 3 substeps, damping 0.997, speed cap 26, radius 0.6 particle spacings.
The body's particles fill a jittered ball instead of the surface extractor's shell.
"""
from __future__ import annotations

import math

import numpy as np

from . import scenes as S
from .batch import FluidBatch
from .edit import SdfEditor, body_centroids_device
from .field import unionField, volumeField
from .pbd import ClaySolver
from .splat import ParticleCloud, SplatField
from .volume import Root, SdfVolume


def body_rest(count: int, radius: float, seed: int) -> np.ndarray:
    """`count` points of a jittered sphere: the cubic lattice points inside a ball of
    `radius` (spacing chosen so that at least `count` fit, nearest to the centre
    first), each jittered by up to a fifth of the spacing with the reference hash."""
    spacing = radius * (4.0 / 3.0 * math.pi / count) ** (1.0 / 3.0)
    while True:
        n = int(math.ceil(2 * radius / spacing)) + 1
        g = (np.arange(n) - (n - 1) / 2) * spacing
        zz, yy, xx = np.meshgrid(g, g, g, indexing="ij")
        p = np.stack([xx.ravel(), yy.ravel(), zz.ravel()], axis=-1)
        r = np.linalg.norm(p, axis=1)
        inside = p[r <= radius]
        if inside.shape[0] >= count:
            break
        spacing *= 0.98
    order = np.argsort(np.linalg.norm(inside, axis=1), kind="stable")[:count]
    pts = inside[order]
    ids = np.arange(count, dtype=np.uint64)
    seeds = ((np.uint64(seed) * np.uint64(9781) + ids) & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    u = S.hash3(seeds).astype(np.float64)
    return (pts + (u * 2 - 1) * spacing * 0.2).astype(np.float32)


FRAME = 1.0 / 60.0
SUBSTEPS = 3


class CoupledBatch:
    def __init__(self, root: Root, fluid_counts, fluid_pos: np.ndarray, fluid_options: dict, world: dict,
                 world_brushes, body_count: int = 4096, body_radius: float = 0.95, body_box: float = 2.7,
                 body_res: int = 32, body_mips: int = 2, body_splat_radius: float = 0.08,
                 erosion=(2.0, 0.12, 0.03), seed: int = 11, orbit: float = 1.5, height: float = 1.0,
                 phases=None, body_seeds=None):
        """Body s starts at (orbit cos phase_s, height, orbit sin phase_s) and falls."""
        import torch

        self.root = root
        counts = np.asarray(fluid_counts, dtype=np.uint32)
        self.scenes = n = len(counts)
        self.body_count = body_count
        self.erosion = erosion
        # world: one instance per scene, baked once and replicated
        self.world = SdfVolume(root, dict(world, instances=n))
        self.world.bakeBrushes(world_brushes)
        self.world.replicate()
        self.world_field = volumeField(self.world)
        # bodies: one solver, body s owns particles [s * body_count, (s + 1) * body_count)
        seeds = body_seeds if body_seeds is not None else [seed + s for s in range(n)]
        rest = np.concatenate([body_rest(body_count, body_radius, int(seeds[s])) for s in range(n)], axis=0)
        ph = np.asarray(phases if phases is not None else np.arange(n) * (2 * math.pi / max(n, 1)), np.float64)
        centre = np.stack([orbit * np.cos(ph), np.full(n, height), orbit * np.sin(ph)], axis=-1)
        start = (rest + np.repeat(centre, body_count, axis=0)).astype(np.float32)
        spacing = body_radius * (4.0 / 3.0 * math.pi / body_count) ** (1.0 / 3.0)
        self.solver = ClaySolver(root, self.world_field,
                                 dict(gravity=(0.0, -18.0, 0.0), dt=FRAME, substeps=SUBSTEPS, radius=0.6 * spacing,
                                      friction=0.45, damping=0.997, maxSpeed=26.0),
                                 capacity=n * body_count, maxBodies=n)
        plasticity = 1.0 - (1.0 - 0.35) ** (FRAME / SUBSTEPS)
        self.solver.setBodies([dict(first=s * body_count, count=body_count, stiffness=0.55, plasticity=plasticity)
                               for s in range(n)])
        self.solver.set_particles(start, rest, np.repeat(np.arange(n, dtype=np.uint32), body_count))
        self.pos = self.solver.positions()
        fc = np.stack([np.arange(n) * body_count, np.full(n, body_count)], axis=-1).astype(np.int32)
        self.body_fc = torch.from_numpy(fc.reshape(-1)).cuda()
        self.body_com = torch.zeros(6 * n, dtype=torch.float64, device="cuda")
        self._centroids(torch.cuda.current_stream())
        self.body_cloud = ParticleCloud(n * body_count, self.pos, live=("prefix", n * body_count))
        self.body = SplatField(root, self.body_cloud,
                               dict(radius=body_splat_radius, resolution=body_res, mipLevels=body_mips, band=4,
                                    worldSize=body_box, origin=(0, 0, 0), mode="tiled"),
                               scenes=[body_count] * n, follow=self.body_com)
        self.collider = unionField(self.world_field, self.body.field)
        self.fluid = FluidBatch(root, self.collider, fluid_options, counts)
        self.fluid.set_state(fluid_pos)
        self.editor = SdfEditor(self.world)

    def _centroids(self, stream) -> None:
        body_centroids_device(self.root, self.pos, self.body_fc, self.body_com, precision="f64", stream=stream)

    def body_positions(self) -> np.ndarray:
        """(scenes * body_count, 3) float32 body particle positions (synchronises)."""
        return self.solver.get("pos")

    def step(self, stream) -> None:
        self.solver.step(stream)
        self._centroids(stream)
        self.body.bake(stream)
        self.fluid.step(stream)
        v, r, b = self.erosion
        self.editor.erodeBatch(self.fluid, v, r, b, stream=stream)

    def destroy(self) -> None:
        self.editor.destroy()
        self.fluid.destroy()
        self.collider.destroy()
        self.body.destroy()
        self.solver.destroy()
        self.world_field.destroy()
        self.world.destroy()


def coupled_scene(scenes: int = 8, fluid_count: int = 524288, world_res: int = 288, seed: int = 5):
    """The coupled workload: per scene a 128 x 32 x 128 dam-break block (524 288
    particles at 0.05, beside the ridge: x in [-3.2, 3.2], z in [1, 7.4]) over the
    ridge world (288^3 x 3 mips over 24 units), one 4096-particle body of radius
    0.95 circling the ridge at height 1, erosion of the world by contacts faster
    than 2 units/s.  Smaller `fluid_count` keeps the block's footprint and thins
    its height."""
    nx = nz = 128
    ny = max(1, fluid_count // (nx * nz))
    block = S.lattice(nx, ny, nz, 0.05, (-3.2, 0.3, 1.0), seed=seed)[:fluid_count]
    pos = np.concatenate([block] * scenes, axis=0)
    opts = S.dam_break_options(1, 64, 3)
    opts.update({"followSize": 6.4, "splatMode": "tiled"})
    world = {"resolution": world_res, "mipLevels": 3, "worldSize": S.WORLD, "origin": S.ORIGIN, "band": 4}
    return {
        "name": f"coupled {scenes} scenes x ({block.shape[0]} fluid + 4096 body), {world_res}^3 world per scene",
        "counts": np.full(scenes, block.shape[0], dtype=np.uint32),
        "pos": pos,
        "options": opts,
        "world": world,
        "world_brushes": S.ridge_world_brushes(),
    }
