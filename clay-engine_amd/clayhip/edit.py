"""SdfEditor (lib/src/field/modify.ts:41-238) and body centroids over the HIP library.

SdfEditor applies brush edits in place to the mip-0 tiles they reach, walks the
dirty-tile mask down the mip chain (TileGrid, tilegrid.ts) and rebuilds each
coarser level over its live tiles (MipRefiner, mips.ts).  `erode` turns a
fluid's contact list into cut spheres on the device and flushes them, with no
host readback in between.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .brush import make_brush
from .volume import SdfVolume


class SdfEditor:
    def __init__(self, volume: SdfVolume, options: dict | None = None):
        o = options or {}
        d = L.EditorDesc(int(o.get("maxEdits", 64)), int(o.get("tileCapacity", 8)), float(o.get("safety", 0.98)))
        h = C.c_void_p()
        L.check(L.lib().clay_editor_create(volume.handle, C.byref(d), C.byref(h)))
        self._h = h
        self.volume = volume
        self.maxEdits = d.max_edits

    def push(self, edit, scene: int = 0) -> bool:
        """SdfEditor.push (modify.ts:200-212): a BrushDesc dict or a Brush record,
        queued for one scene of a multi-instance volume (0 otherwise)."""
        b = edit if isinstance(edit, L.Brush) else make_brush(edit)
        accepted = C.c_int32()
        L.check(L.lib().clay_editor_push_scene(self._h, int(scene), C.byref(b), 1, C.byref(accepted)))
        return accepted.value == 1

    @property
    def pendingCount(self) -> int:
        n = C.c_int32()
        L.check(L.lib().clay_editor_pending(self._h, C.byref(n)))
        return n.value

    def flush(self, stream=None) -> bool:
        """modify.ts:214-238; False when nothing was queued."""
        done = C.c_int32()
        L.check(L.lib().clay_editor_flush(self._h, L.stream_handle(stream), C.byref(done)))
        return done.value == 1

    def erode(self, fluid, min_speed: float, radius: float, blend: float = 0.02, scene: int | None = None,
              stream=None) -> None:
        """Cut spheres at the fluid's fast contacts, flushed on the device."""
        if scene is None:
            L.check(L.lib().clay_editor_erode(self._h, fluid.handle, float(min_speed), float(radius), float(blend),
                                              L.stream_handle(stream)))
        else:
            L.check(L.lib().clay_editor_erode_scene(self._h, fluid._h, int(scene), float(min_speed), float(radius),
                                                    float(blend), L.stream_handle(stream)))

    def erodeBatch(self, batch, min_speed: float, radius: float, blend: float = 0.02, stream=None) -> None:
        """Every scene's contacts -> that scene's world instance (one instance per scene)."""
        L.check(L.lib().clay_editor_erode_batch(self._h, batch._h, float(min_speed), float(radius), float(blend),
                                                L.stream_handle(stream)))

    def refreshAll(self, stream=None) -> None:
        L.check(L.lib().clay_editor_refresh_all(self._h, L.stream_handle(stream)))

    def readTileCounts(self, scene: int = 0) -> list[int]:
        """TileGrid.readTileCounts (tilegrid.ts:281-297)."""
        out = (C.c_uint32 * 16)()
        L.check(L.lib().clay_editor_tile_counts_scene(self._h, int(scene), out))
        return [out[m] for m in range(self.volume.mipLevels)]

    def cellCounts(self, scene: int = 0) -> np.ndarray:
        out = np.empty(self.volume.totalCells, dtype=np.uint32)
        L.check(L.lib().clay_editor_cells_download_scene(self._h, int(scene), out.ctypes.data))
        return out

    def destroy(self) -> None:
        if self._h is not None:
            L.lib().clay_editor_destroy(self._h)
        self._h = None


def body_centroids(root, pos, first_count, rest=None, precision: str = "f64", stream=None) -> np.ndarray:
    """ClaySolver.match's centroids (pbd.ts:240-266) of bodies [first, first + count)
    of a device particle set.  pos / rest: three device float arrays each (torch
    tensors or int pointers).  Returns (bodies, 6) float64: com, restCom."""
    def ptrs(arrs):
        return (C.c_void_p * 3)(*[a if isinstance(a, int) else a.data_ptr() for a in arrs])

    fc = np.ascontiguousarray(np.asarray(first_count, dtype=np.uint32).reshape(-1, 2))
    out = np.zeros((fc.shape[0], 6), dtype=np.float64)
    mode = {"f32": L.REDUCE_F32_TREE, "f64": L.REDUCE_F64}[precision]
    L.check(L.lib().clay_body_centroids(root.handle, ptrs(pos), None if rest is None else ptrs(rest),
                                        fc.ctypes.data, fc.shape[0], mode, out.ctypes.data, L.stream_handle(stream)))
    return out


def body_centroids_device(root, pos, first_count_dev, out_dev, rest=None, precision: str = "f64", stream=None) -> None:
    """Enqueues the centroids of bodies into `out_dev` (CUDA float64, 6 per body)
    from a CUDA int32/uint32 table of (first, count) pairs; nothing is read back."""
    def ptrs(arrs):
        return (C.c_void_p * 3)(*[a if isinstance(a, int) else a.data_ptr() for a in arrs])

    mode = {"f32": L.REDUCE_F32_TREE, "f64": L.REDUCE_F64}[precision]
    n = first_count_dev.numel() // 2
    L.check(L.lib().clay_body_centroids_device(root.handle, ptrs(pos), None if rest is None else ptrs(rest),
                                               C.c_void_p(first_count_dev.data_ptr()), n, mode,
                                               C.c_void_p(out_dev.data_ptr()), L.stream_handle(stream)))
