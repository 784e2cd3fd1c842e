"""ClaySolver, ParticleSet and SurfaceExtractor (lib/src/sim/pbd.ts:78-424,
particles.ts:105-206, extract.ts:85-484) over the HIP library.

A ParticleSet's particle and body records live on the GPU.  Bodies are declared
with `addBody` (the reference's defaults: stiffness 0.5, plasticity 0, flags 0)
and own slices of particles, which a SurfaceExtractor fills from a field (or
`set_particles` writes directly).  `step` enqueues one ClaySolver.step: the kick
when an impulse is pending, then per substep predict, match, apply and finalize.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .volume import Root

FIELDS = {"pos": 0, "prev": 1, "vel": 2, "velPrev": 3, "rest": 4, "restNormal": 5}


def solver_desc(**o) -> L.SolverDesc:
    d = L.SolverDesc()
    L.lib().clay_solver_desc_default(C.byref(d))
    if "gravity" in o:
        for i in range(3):
            d.gravity[i] = o["gravity"][i]
    for key, field in (("dt", "dt"), ("substeps", "substeps"), ("rotationIterations", "rotation_iterations"),
                       ("radius", "radius"), ("friction", "friction"), ("damping", "damping"),
                       ("maxSpeed", "max_speed")):
        if key in o:
            setattr(d, field, o[key])
    return d


class ClaySolver:
    """new ParticleSet(root, {capacity, maxBodies}) + new ClaySolver(set, world, options)."""

    def __init__(self, root: Root, world, options: dict | None = None, capacity: int = 1 << 16,
                 maxBodies: int = 32):
        d = solver_desc(**(options or {}))
        h = C.c_void_p()
        wh = world.handle if world is not None else None
        L.check(L.lib().clay_solver_create(root.handle, wh, C.byref(d), int(capacity), int(maxBodies), C.byref(h)))
        self._h = h
        self.root, self.world = root, world
        self.capacity, self.maxBodies = int(capacity), int(maxBodies)
        self.substeps, self.dt = d.substeps, d.dt
        self._bodies: list[dict] = []

    @property
    def handle(self):
        return self._h

    @property
    def bodyCount(self) -> int:
        return len(self._bodies)

    def addBody(self, first: int = 0, count: int = 0, stiffness: float = 0.5, plasticity: float = 0.0,
                flags: int = 0) -> int:
        """ParticleSet.addBody (particles.ts:146-168).  The slice may be given here, or
        left empty for a SurfaceExtractor to fill in."""
        d = L.BodyDesc(int(first), int(count), int(flags), float(stiffness), float(plasticity), 0.0)
        idx = C.c_uint32()
        L.check(L.lib().clay_solver_add_body(self._h, C.byref(d), C.byref(idx)))
        self._bodies.append(dict(first=first, count=count, stiffness=stiffness, plasticity=plasticity, flags=flags))
        return idx.value

    def setBodies(self, bodies) -> None:
        """Declares every body at once: dicts of first, count and optionally
        stiffness, plasticity, flags (addBody's defaults otherwise)."""
        if len(bodies) > self.maxBodies:
            raise L.ClayError(L.ERR_INVALID, f"ParticleSet: body limit {self.maxBodies} reached")
        self._bodies = [dict(first=int(b["first"]), count=int(b["count"]), stiffness=b.get("stiffness", 0.5),
                             plasticity=b.get("plasticity", 0.0), flags=b.get("flags", 0)) for b in bodies]
        self._upload_bodies()

    def _upload_bodies(self) -> None:
        arr = (L.BodyDesc * max(1, len(self._bodies)))()
        for i, b in enumerate(self._bodies):
            arr[i] = L.BodyDesc(b["first"], b["count"], b["flags"], b["stiffness"], b["plasticity"], 0.0)
        L.check(L.lib().clay_solver_set_bodies(self._h, arr, len(self._bodies)))

    def set_particles(self, pos: np.ndarray, rest: np.ndarray, body: np.ndarray) -> None:
        """Particle slots [0, n): positions and rest positions (n, 3), owners (n,)."""
        p = np.ascontiguousarray(np.asarray(pos, dtype=np.float32).T)
        r = np.ascontiguousarray(np.asarray(rest, dtype=np.float32).T)
        b = np.ascontiguousarray(np.asarray(body, dtype=np.uint32))
        L.check(L.lib().clay_solver_set_particles(self._h, p.ctypes.data, r.ctypes.data, b.ctypes.data, p.shape[1]))

    def setForce(self, body: int, accel, impulse=(0.0, 0.0, 0.0)) -> None:
        """ClaySolver.setForce (pbd.ts:383-394)."""
        a = (C.c_double * 3)(*accel)
        i = (C.c_double * 3)(*impulse)
        L.check(L.lib().clay_solver_set_force(self._h, int(body), a, i))

    def step(self, stream=None) -> None:
        """ClaySolver.step (pbd.ts:400-415)."""
        L.check(L.lib().clay_solver_step(self._h, L.stream_handle(stream)))

    def get(self, what: str) -> np.ndarray:
        """(capacity, 3) float32 of pos / prev / vel / velPrev / rest / restNormal, or
        (capacity,) of material."""
        if what == "material":
            out = np.empty(self.capacity, dtype=np.float32)
            L.check(L.lib().clay_solver_get_particles(self._h, 6, out.ctypes.data))
            return out
        out = np.empty((3, self.capacity), dtype=np.float32)
        L.check(L.lib().clay_solver_get_particles(self._h, FIELDS[what], out.ctypes.data))
        return out.T.copy()

    def owners(self) -> np.ndarray:
        out = np.empty(self.capacity, dtype=np.uint32)
        L.check(L.lib().clay_solver_get_owners(self._h, out.ctypes.data))
        return out

    def bodies(self) -> np.ndarray:
        """(maxBodies, 10): com, restCom, rot xyzw."""
        out = np.empty((self.maxBodies, 10), dtype=np.float32)
        L.check(L.lib().clay_solver_get_bodies(self._h, out.ctypes.data))
        return out

    def slices(self) -> np.ndarray:
        """(maxBodies, 3) u32: first, count, flags."""
        out = np.empty((self.maxBodies, 3), dtype=np.uint32)
        L.check(L.lib().clay_solver_get_slices(self._h, out.ctypes.data))
        return out

    def positions(self) -> tuple[int, int, int]:
        """Device addresses of the x, y, z position planes."""
        arr = (C.c_void_p * 3)()
        L.check(L.lib().clay_solver_positions(self._h, arr))
        return tuple(int(a) for a in arr)

    def destroy(self) -> None:
        if self._h:
            L.lib().clay_solver_destroy(self._h)
            self._h = None


class SurfaceExtractor:
    """new SurfaceExtractor(set, field, {resolution, body, base, capacity, motion})
    (extract.ts:103-444) over a ClaySolver's particle set.  Particle slots are
    numbered in ascending cell index and quads in ascending lattice point then edge
    axis, one of the orders the reference's atomic counters allow."""

    def __init__(self, solver: ClaySolver, field, options: dict):
        d = L.ExtractorDesc()
        d.resolution = int(options.get("resolution", 32))
        d.body = int(options["body"])
        d.base = int(options["base"])
        d.capacity = int(options["capacity"])
        d.motion = 1 if options.get("motion") else 0
        d.instance = int(options.get("instance", 0))
        h = C.c_void_p()
        L.check(L.lib().clay_extractor_create(solver.handle, field.handle, C.byref(d), C.byref(h)))
        self._h = h
        self.solver, self.field = solver, field
        self.resolution, self.body, self.base, self.capacity = d.resolution, d.body, d.base, d.capacity
        q = C.c_uint32()
        L.check(L.lib().clay_extractor_quad_capacity(self._h, C.byref(q)))
        self.quadCapacity = q.value

    def setRegion(self, origin, size: float, vel=(0.0, 0.0, 0.0), preserveMotion: bool = False) -> None:
        """SurfaceExtractor.setRegion (extract.ts:451-463)."""
        o = (C.c_double * 3)(*origin)
        v = (C.c_double * 3)(*vel)
        L.check(L.lib().clay_extractor_set_region(self._h, o, float(size), v, 1 if preserveMotion else 0))

    def extract(self, stream=None) -> None:
        """SurfaceExtractor.extract (extract.ts:469-475)."""
        L.check(L.lib().clay_extractor_extract(self._h, L.stream_handle(stream)))

    def read(self):
        """(counts (particles, quads found), draw args (4,), indices (6 * quadCapacity,))."""
        counts = np.zeros(2, dtype=np.uint32)
        draw = np.zeros(4, dtype=np.uint32)
        idx = np.zeros(6 * max(self.quadCapacity, 1), dtype=np.uint32)
        L.check(L.lib().clay_extractor_read(self._h, counts.ctypes.data, draw.ctypes.data, idx.ctypes.data))
        return counts, draw, idx

    def motion(self, value=None) -> np.ndarray:
        """The ExtractMotion record (centre xyz, 0, velocity xyz, 0); writes `value` first."""
        out = np.zeros(8, dtype=np.float32)
        src = None if value is None else np.ascontiguousarray(value, dtype=np.float32)
        L.check(L.lib().clay_extractor_motion(self._h, None if src is None else src.ctypes.data, out.ctypes.data))
        return out

    def destroy(self) -> None:
        if self._h:
            L.lib().clay_extractor_destroy(self._h)
            self._h = None
