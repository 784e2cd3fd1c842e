"""ClaySolver and ParticleSet (lib/src/sim/pbd.ts:78-424, particles.ts:105-206) over the HIP library.

A ParticleSet's particle and body records live on the GPU.  Bodies are declared
with `addBody` (the reference's defaults: stiffness 0.5, plasticity 0, flags 0)
and own slices of particles; `set_particles` plays the extractor's part of
writing positions, rest positions and owners.  `step` enqueues one
ClaySolver.step: the kick when an impulse is pending, then per substep predict,
match, apply and finalize.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .volume import Root

FIELDS = {"pos": 0, "prev": 1, "vel": 2, "velPrev": 3, "rest": 4}


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

    def addBody(self, first: int, count: int, stiffness: float = 0.5, plasticity: float = 0.0, flags: int = 0) -> int:
        """ParticleSet.addBody (particles.ts:146-168); the slice is given here instead
        of being filled in by an extractor."""
        if len(self._bodies) >= self.maxBodies:
            raise L.ClayError(L.ERR_INVALID, f"ParticleSet: body limit {self.maxBodies} reached")
        self._bodies.append(dict(first=first, count=count, stiffness=stiffness, plasticity=plasticity, flags=flags))
        self._upload_bodies()
        return len(self._bodies) - 1

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
        """(capacity, 3) float32 of pos / prev / vel / velPrev / rest."""
        out = np.empty((3, self.capacity), dtype=np.float32)
        L.check(L.lib().clay_solver_get_particles(self._h, FIELDS[what], out.ctypes.data))
        return out.T.copy()

    def bodies(self) -> np.ndarray:
        """(maxBodies, 10): com, restCom, rot xyzw."""
        out = np.empty((self.maxBodies, 10), dtype=np.float32)
        L.check(L.lib().clay_solver_get_bodies(self._h, out.ctypes.data))
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
