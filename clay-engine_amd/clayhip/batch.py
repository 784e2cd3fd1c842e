"""Batched independent scenes over the HIP library (clay_batch_*).

The reference runs one FluidSim per scene (lib/src/sim/fluid.ts:111-582), each
with its own dispatches.  FluidBatch steps many of them with the same launches;
every scene's result is the one a FluidSim of the same options computes for that
scene alone.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .field import TracedField
from .fluid import fluid_desc
from .volume import Root

_STATE = {"pos": L.STATE_POS, "vel": L.STATE_VEL, "density": L.STATE_DENSITY,
          "pressure": L.STATE_PRESSURE, "nvel": L.STATE_NVEL}


class FluidBatch:
    def __init__(self, root: Root, world: TracedField | None, options: dict, counts):
        counts = np.ascontiguousarray(np.asarray(counts, dtype=np.uint32))
        self.root = root
        self.world = world
        self.counts = counts
        self.total = int(counts.sum())
        self._counts_c = counts.ctypes.data_as(C.POINTER(C.c_uint32))
        d = L.BatchDesc()
        d.scenes = counts.shape[0]
        d.counts = self._counts_c
        d.fluid = fluid_desc(dict(options, capacity=max(self.total, 1)))
        h = C.c_void_p()
        L.check(L.lib().clay_batch_create(root.handle, world.handle if world else None, C.byref(d), C.byref(h)))
        self._h = h
        k = L.FluidConsts()
        L.check(L.lib().clay_batch_constants(self._h, C.byref(k)))
        self.constants = k
        self.mips = d.fluid.bake_mip_levels
        self.resolution = d.fluid.bake_resolution

    @property
    def scenes(self) -> int:
        return int(self.counts.shape[0])

    def scene(self, i: int) -> tuple[int, int]:
        first, count = C.c_uint32(), C.c_uint32()
        L.check(L.lib().clay_batch_scene(self._h, int(i), C.byref(first), C.byref(count)))
        return first.value, count.value

    def set_state(self, pos: np.ndarray, vel: np.ndarray | None = None) -> None:
        p = np.ascontiguousarray(np.asarray(pos, dtype=np.float32).T)
        if p.shape[1] != self.total:
            raise ValueError(f"FluidBatch.set_state: expected {self.total} particles, got {p.shape[1]}")
        v = None if vel is None else np.ascontiguousarray(np.asarray(vel, dtype=np.float32).T)
        L.check(L.lib().clay_batch_set_state(self._h, p.ctypes.data, None if v is None else v.ctypes.data))

    def get_state(self, what: str) -> np.ndarray:
        n = self.total
        if what in ("density", "pressure"):
            out = np.empty(n, dtype=np.float32)
            L.check(L.lib().clay_batch_get_state(self._h, _STATE[what], out.ctypes.data))
            return out
        out = np.empty((3, n), dtype=np.float32)
        L.check(L.lib().clay_batch_get_state(self._h, _STATE[what], out.ctypes.data))
        return out.T.copy()

    def step(self, stream=None) -> None:
        L.check(L.lib().clay_batch_step(self._h, L.stream_handle(stream)))

    def set_graph(self, enabled: bool) -> None:
        L.check(L.lib().clay_batch_set_graph(self._h, 1 if enabled else 0))

    @property
    def tuning(self) -> dict:
        t = L.Tuning()
        L.check(L.lib().clay_batch_get_tuning(self._h, C.byref(t)))
        return {f: getattr(t, f) for f, _ in L.Tuning._fields_}

    @tuning.setter
    def tuning(self, values: dict) -> None:
        cur = self.tuning
        cur.update(values)
        t = L.Tuning(**cur)
        L.check(L.lib().clay_batch_set_tuning(self._h, C.byref(t)))

    def read_contacts(self, scene: int) -> tuple[int, list]:
        pts = np.zeros((L.CONTACTS, 4), dtype=np.float32)
        n, raw = C.c_uint32(), C.c_uint32()
        L.check(L.lib().clay_batch_read_contacts(self._h, int(scene), pts.ctypes.data, C.byref(n), C.byref(raw)))
        return raw.value, [tuple(float(v) for v in pts[i]) for i in range(n.value)]

    def download(self, scene: int, mip: int) -> np.ndarray:
        r = self.resolution >> mip
        out = np.empty((r, r, r, 4), dtype=np.uint16)
        L.check(L.lib().clay_batch_volume_download(self._h, int(scene), int(mip), out.ctypes.data))
        return out

    def codes(self, scene: int, mip: int) -> np.ndarray:
        r = self.resolution >> mip
        out = np.empty(r ** 3, dtype=np.uint32)
        L.check(L.lib().clay_batch_grid_download(self._h, int(scene), int(mip), out.ctypes.data))
        return out

    def live_tiles(self, scene: int, mip: int) -> np.ndarray:
        tr = (self.resolution >> mip) // 8
        out = np.empty(tr ** 3, dtype=np.uint8)
        L.check(L.lib().clay_batch_tiles_download(self._h, int(scene), int(mip), out.ctypes.data))
        return out

    def origin(self, scene: int) -> np.ndarray:
        out = (C.c_float * 3)()
        L.check(L.lib().clay_batch_origin(self._h, int(scene), out))
        return np.array(out[:], dtype=np.float32)

    def conserved(self, scene: int) -> dict:
        out = (C.c_double * 7)()
        L.check(L.lib().clay_batch_conserved(self._h, int(scene), out))
        return {"mass": out[0], "momentum": tuple(out[1:4]), "com": tuple(out[4:7])}

    def splat_atomics(self) -> tuple[int, int]:
        """Bakes every scene's surface of the current state again and counts its
        owner walk's atomics: (LDS ds_min_u32, global atomicMin)."""
        out = (C.c_uint64 * 2)()
        L.check(L.lib().clay_batch_splat_atomics(self._h, out))
        return int(out[0]), int(out[1])

    def destroy(self) -> None:
        if self._h is not None:
            L.lib().clay_batch_destroy(self._h)
        self._h = None
