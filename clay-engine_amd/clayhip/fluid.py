"""FluidSim (lib/src/sim/fluid.ts:111-582) over the HIP library."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .field import TracedField
from .splat import SplatField
from .volume import Root

# FluidSimOptions key -> clay_fluid_desc field (fluid.ts:60-91)
_OPTION_FIELDS = {
    "capacity": "capacity",
    "bucket": "bucket",
    "spacing": "spacing",
    "smoothing": "smoothing",
    "density": "density",
    "radius": "radius",
    "surfaceRadius": "surface_radius",
    "material": "material",
    "bakeResolution": "bake_resolution",
    "bakeMipLevels": "bake_mip_levels",
    "bakeBand": "bake_band",
    "worldSize": "world_size",
    "substeps": "substeps",
    "stiffness": "stiffness",
    "viscosity": "viscosity",
    "damping": "damping",
    "maxSpeed": "max_speed",
    "followSize": "follow_size",
}
# FluidParams members configure() may patch (fluid.ts:528-530)
_PARAM_FIELDS = {
    "gravity": "gravity",
    "dt": "dt",
    "emitPos": "emit_pos",
    "emitRadius": "emit_radius",
    "emitVel": "emit_vel",
    "maxSpeed": "max_speed",
    "stiffness": "stiffness",
    "viscosity": "viscosity",
    "damping": "damping",
    "contactDepth": "contact_depth",
}
_FORBIDDEN = {"count", "emitBase", "emitCount", "emitSeed"}


def fluid_desc(options: dict | None = None) -> L.FluidDesc:
    """clay_fluid_desc with the reference defaults, patched by FluidSimOptions."""
    d = L.FluidDesc()
    L.lib().clay_fluid_desc_default(C.byref(d))
    for key, value in (options or {}).items():
        if key in _OPTION_FIELDS:
            setattr(d, _OPTION_FIELDS[key], value)
        elif key in ("origin", "gravity"):
            arr = getattr(d, key)
            for i in range(3):
                arr[i] = float(value[i])
        elif key == "splatMode":
            d.splat_mode = {"plain": L.SPLAT_PLAIN, "tiled": L.SPLAT_TILED}[value]
        elif key == "precision":
            d.precision = {"f32": L.PRECISION_F32, "fp16lds": L.PRECISION_FP16LDS}[value]
        elif key == "label":
            pass
        else:
            raise ValueError(f"FluidSim: unknown option {key!r}")
    return d


class FluidSim:
    """Weakly compressible SPH with XSPH viscosity colliding against a TracedField,
    baked every step into a SplatField (fluid.ts:93-110)."""

    def __init__(self, root: Root, world: TracedField | None, options: dict | None = None, **kw):
        opts = dict(options or {}, **kw)
        self.root = root
        self.world = world
        d = fluid_desc(opts)
        h = C.c_void_p()
        L.check(L.lib().clay_fluid_create(root.handle, world.handle if world else None, C.byref(d), C.byref(h)))
        self._h = h
        self.capacity = d.capacity
        self.spacing = d.spacing
        sh = C.c_void_p()
        L.check(L.lib().clay_fluid_surface(self._h, C.byref(sh)))
        self.surface = SplatField(root, None, _handle=sh)
        self.field = self.surface.field
        k = L.FluidConsts()
        L.check(L.lib().clay_fluid_constants(self._h, C.byref(k)))
        self.constants = k
        self.substeps = k.substeps

    @property
    def handle(self):
        return self._h

    @property
    def particleCount(self) -> int:
        out = C.c_uint32()
        L.check(L.lib().clay_fluid_particle_count(self._h, C.byref(out)))
        return out.value

    @property
    def params(self) -> dict:
        p = L.FluidParams()
        L.check(L.lib().clay_fluid_get_params(self._h, C.byref(p)))
        out = {}
        for key, field in _PARAM_FIELDS.items():
            v = getattr(p, field)
            out[key] = tuple(v) if hasattr(v, "__len__") else v
        return out

    def spawn(self, n: int, pos, vel, spread: float | None = None) -> None:
        """fluid.ts:511-525."""
        L.check(L.lib().clay_fluid_spawn(self._h, int(n), L.dvec(pos), L.dvec(vel),
                                         -1.0 if spread is None else float(spread)))

    def configure(self, patch: dict) -> None:
        """fluid.ts:528-530: anything in FluidParams except the emission cursor."""
        bad = set(patch) & _FORBIDDEN
        if bad:
            raise ValueError(f"FluidSim.configure: {sorted(bad)} are managed by the simulation")
        p = L.FluidParams()
        L.check(L.lib().clay_fluid_get_params(self._h, C.byref(p)))
        for key, value in patch.items():
            if key not in _PARAM_FIELDS:
                raise ValueError(f"FluidSim.configure: unknown parameter {key!r}")
            field = _PARAM_FIELDS[key]
            cur = getattr(p, field)
            if hasattr(cur, "__len__"):
                for i in range(3):
                    cur[i] = float(value[i])
            else:
                setattr(p, field, float(value))
        L.check(L.lib().clay_fluid_configure(self._h, C.byref(p)))

    def step(self, stream=None) -> None:
        """fluid.ts:536-558: emit, `substeps` substeps, bake every mip."""
        L.check(L.lib().clay_fluid_step(self._h, L.stream_handle(stream)))

    def set_graph(self, enabled: bool) -> None:
        """Replay each frame from a captured HIP graph (no reference counterpart)."""
        L.check(L.lib().clay_fluid_set_graph(self._h, 1 if enabled else 0))

    def substep(self, stream=None) -> None:
        L.check(L.lib().clay_fluid_substep(self._h, L.stream_handle(stream)))

    def readContacts(self) -> list[tuple[float, float, float, float]]:
        """fluid.ts:565-569 (synchronous here: the stream is drained first)."""
        pts = np.zeros((L.CONTACTS, 4), dtype=np.float32)
        n = C.c_uint32()
        raw = C.c_uint32()
        L.check(L.lib().clay_fluid_read_contacts(self._h, pts.ctypes.data, C.byref(n), C.byref(raw)))
        self.lastContactCount = raw.value
        return [tuple(float(v) for v in pts[i]) for i in range(n.value)]

    @property
    def tuning(self) -> dict:
        t = L.Tuning()
        L.check(L.lib().clay_fluid_get_tuning(self._h, C.byref(t)))
        return {f: getattr(t, f) for f, _ in L.Tuning._fields_}

    @tuning.setter
    def tuning(self, values: dict) -> None:
        cur = self.tuning
        cur.update(values)
        t = L.Tuning(**cur)
        L.check(L.lib().clay_fluid_set_tuning(self._h, C.byref(t)))

    def splat_atomics(self) -> tuple[int, int]:
        """Bakes the surface of the current state again and counts its owner walk's
        atomics: (LDS ds_min_u32, global atomicMin)."""
        out = (C.c_uint64 * 2)()
        L.check(L.lib().clay_fluid_splat_atomics(self._h, out))
        return int(out[0]), int(out[1])

    def conserved(self) -> dict:
        """Total mass, momentum and centre of mass of the live particles, reduced in FP64."""
        out = (C.c_double * 7)()
        L.check(L.lib().clay_fluid_conserved(self._h, out))
        return {"mass": out[0], "momentum": tuple(out[1:4]), "com": tuple(out[4:7])}

    # ---- state (the build's seeded-lattice entry and readback) ----
    def set_state(self, pos: np.ndarray, vel: np.ndarray | None = None) -> None:
        """pos/vel as (n, 3) float32; the live count becomes n."""
        p = np.ascontiguousarray(np.asarray(pos, dtype=np.float32).T)
        v = None if vel is None else np.ascontiguousarray(np.asarray(vel, dtype=np.float32).T)
        n = p.shape[1]
        L.check(L.lib().clay_fluid_set_state(self._h, p.ctypes.data, None if v is None else v.ctypes.data, n))

    def get_state(self, what: str) -> np.ndarray:
        idx = {"pos": L.STATE_POS, "vel": L.STATE_VEL, "density": L.STATE_DENSITY,
               "pressure": L.STATE_PRESSURE, "nvel": L.STATE_NVEL}[what]
        n = self.particleCount
        if what in ("density", "pressure"):
            out = np.empty(n, dtype=np.float32)
            L.check(L.lib().clay_fluid_get_state(self._h, idx, out.ctypes.data))
            return out
        out = np.empty((3, n), dtype=np.float32)
        L.check(L.lib().clay_fluid_get_state(self._h, idx, out.ctypes.data))
        return out.T.copy()

    def debug(self, what: str) -> np.ndarray:
        """Sorted cell keys ('keys') or the sort permutation ('perm') of the last substep."""
        n = self.particleCount
        out = np.empty(n, dtype=np.uint32)
        L.check(L.lib().clay_fluid_debug_download(self._h, {"keys": L.DEBUG_KEYS, "perm": L.DEBUG_PERM}[what],
                                                  out.ctypes.data))
        return out

    def destroy(self) -> None:
        if self._h is not None:
            L.lib().clay_fluid_destroy(self._h)
        self._h = None
