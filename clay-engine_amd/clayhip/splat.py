"""ParticleCloud and SplatField (lib/src/sim/splat.ts)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .field import TracedField
from .volume import Root, SdfVolume


def _ptr(a) -> int:
    """Device address of a torch CUDA tensor, or an int taken as an address."""
    if a is None:
        return 0
    if isinstance(a, int):
        return a
    if hasattr(a, "data_ptr"):
        if not getattr(a, "is_cuda", False):
            raise ValueError("ParticleCloud: arrays must live on the GPU")
        if a.dtype.__str__() != "torch.float32" or not a.is_contiguous():
            raise ValueError("ParticleCloud: arrays must be contiguous float32")
        return int(a.data_ptr())
    raise TypeError(f"not a device array: {type(a)!r}")


class ParticleCloud:
    """ParticleCloud (splat.ts:15-38) as device SoA arrays and a live rule:
    None/'all', ('prefix', count) or ('range', first, count)."""

    def __init__(self, capacity: int, positions, live=None, normals=None):
        if len(positions) != 3:
            raise ValueError("ParticleCloud: positions are three arrays (x, y, z)")
        self.capacity = int(capacity)
        self._arrays = (tuple(positions), tuple(normals) if normals else None)
        d = L.CloudDesc()
        d.capacity = self.capacity
        rule, first, count = L.LIVE_ALL, 0, self.capacity
        if live is not None and live != "all":
            if live[0] == "prefix":
                rule, count = L.LIVE_PREFIX, int(live[1])
            elif live[0] == "range":
                rule, first, count = L.LIVE_RANGE, int(live[1]), int(live[2])
            else:
                raise ValueError(f"ParticleCloud: unknown live rule {live!r}")
        d.live_rule, d.live_first, d.live_count = rule, first, count
        for i in range(3):
            d.pos[i] = _ptr(positions[i])
            d.nrm[i] = _ptr(normals[i]) if normals else 0
        self.desc = d


def splat_desc(opts: dict) -> L.SplatDesc:
    """SplatFieldOptions (splat.ts:80-106) as the C descriptor."""
    if "radius" not in opts:
        raise ValueError("SplatField: `radius` is required")
    d = L.SplatDesc()
    d.radius = float(opts["radius"])
    d.material = float(opts.get("material", 0))
    d.resolution = int(opts.get("resolution", 64))
    d.mip_levels = int(opts.get("mipLevels", 0) or 0)
    d.band = float(opts.get("band", 2))
    origin = opts.get("origin", (-12, -4, -12))
    for i in range(3):
        d.origin[i] = float(origin[i])
    d.world_size = float(opts.get("worldSize", 24))
    mode = opts.get("mode", "plain")
    d.mode = {"plain": L.SPLAT_PLAIN, "tiled": L.SPLAT_TILED}[mode]
    d.keep_codes = 1 if opts.get("keepCodes", False) else 0
    return d


def bake_static(root: Root, cloud: ParticleCloud, options: dict, dist, stream=None) -> None:
    """The plain bake of `cloud` with no allocation (clay_splat_bake_static): writes
    every mip's f16 distance plane, concatenated mip-major and x fastest, into `dist`,
    a contiguous CUDA int16/uint16 tensor of sum_m (res >> m)^3 elements."""
    d = splat_desc(dict(options, mode="plain"))
    res = d.resolution if d.resolution > 0 else 64
    mips = d.mip_levels if d.mip_levels > 0 else max(1, int(np.log2(res / 8)))
    need = sum((res >> m) ** 3 for m in range(mips))
    if not getattr(dist, "is_cuda", False) or dist.element_size() != 2 or not dist.is_contiguous():
        raise ValueError("bake_static: dist must be a contiguous 16-bit CUDA tensor")
    if dist.numel() < need:
        raise ValueError(f"bake_static: dist holds {dist.numel()} texels, the bake writes {need}")
    L.check(L.lib().clay_splat_bake_static(root.handle, C.byref(cloud.desc), C.byref(d), C.c_void_p(int(dist.data_ptr())),
                                           L.stream_handle(stream)))


class SplatField:
    """Turns a ParticleCloud into a mip-mapped volume by a fixed-point atomicMin
    union of spheres (splat.ts:136-309).  `mode` 'plain' keeps the reference's
    dense schedule; 'tiled' bins particles to LDS bricks and writes live tiles only.
    Both give the same u32 codes."""

    def __init__(self, root: Root, cloud: ParticleCloud | None, options: dict | None = None, *,
                 _handle=None, scenes=None, follow=None, **kw):
        """`scenes` (particle counts, scene-major) bakes scene s's particles into
        volume instance s; `follow` is None, 'bounds', or a CUDA float64 tensor of 6
        values per scene (xyz first: body centroids) each instance centres on."""
        self.root = root
        self.cloud = cloud
        self._keep = []
        if _handle is not None:
            self._h = _handle
            self._owned = False
        else:
            d = splat_desc(dict(options or {}, **kw))
            h = C.c_void_p()
            if scenes is None:
                L.check(L.lib().clay_splat_create(root.handle, C.byref(cloud.desc), C.byref(d), C.byref(h)))
            else:
                counts = np.ascontiguousarray(np.asarray(scenes, dtype=np.uint32))
                if follow is None:
                    mode, centers = L.FOLLOW_NONE, None
                elif isinstance(follow, str) and follow == "bounds":
                    mode, centers = L.FOLLOW_BOUNDS, None
                else:
                    if str(follow.dtype) != "torch.float64" or not follow.is_cuda:
                        raise ValueError("SplatField: follow centres are a CUDA float64 tensor")
                    mode, centers = L.FOLLOW_CENTERS, C.c_void_p(int(follow.data_ptr()))
                    self._keep.append(follow)
                L.check(L.lib().clay_splat_create_scenes(root.handle, C.byref(cloud.desc), C.byref(d), counts.shape[0],
                                                         counts.ctypes.data, mode, centers, C.byref(h)))
            self.scenes = None if scenes is None else len(scenes)
            self._h = h
            self._owned = True
        vh = C.c_void_p()
        L.check(L.lib().clay_splat_volume(self._h, C.byref(vh)))
        self.volume = SdfVolume(root, _handle=vh)
        fh = C.c_void_p()
        L.check(L.lib().clay_splat_field(self._h, C.byref(fh)))
        self.field = TracedField(root, fh, owned=False, keep=(self,))

    @property
    def handle(self):
        return self._h

    def setLive(self, first: int, count: int) -> None:
        L.check(L.lib().clay_splat_set_live(self._h, int(first), int(count)))

    def bake(self, stream=None) -> None:
        """splat.ts:296-304."""
        L.check(L.lib().clay_splat_bake(self._h, L.stream_handle(stream)))

    def setCenter(self, center, stream=None) -> None:
        """splat.ts:284-287."""
        L.check(L.lib().clay_splat_set_center(self._h, L.dvec(center), L.stream_handle(stream)))
        self.volume._load_info()

    def setSchedule(self, chunk: int = 1024, groups_per_cu: int = 8, xcd: bool = False, owner: bool = True) -> None:
        """Launch shape of the tiled bake (results do not depend on it)."""
        t = L.Tuning(block=256, stage=128, chunk=int(chunk), groups_per_cu=int(groups_per_cu), fluid_xcd=0,
                     splat_xcd=1 if xcd else 0, splat_owner=1 if owner else 0, exact_lds=1)
        L.check(L.lib().clay_splat_set_tuning(self._h, C.byref(t)))

    def codes(self, mip: int = 0) -> np.ndarray:
        """The u32 code grid of one mip, shape (r, r, r) indexed [z, y, x]."""
        r = self.volume.resolution >> mip
        out = np.empty((r, r, r), dtype=np.uint32)
        L.check(L.lib().clay_splat_grid_download(self._h, mip, out.ctypes.data))
        return out

    def sceneCodes(self, scene: int, mip: int = 0) -> np.ndarray:
        """One scene's u32 codes of a scene-partitioned splat, shape (r, r, r)."""
        r = self.volume.resolution >> mip
        out = np.empty((r, r, r), dtype=np.uint32)
        L.check(L.lib().clay_splat_instance_download(self._h, int(scene), mip, out.ctypes.data))
        return out

    def liveTiles(self, mip: int = 0) -> np.ndarray:
        """Exact live-tile mask (some voxel < 65535) of one mip, shape (t, t, t)."""
        t = (self.volume.resolution >> mip) // 8
        out = np.empty((t, t, t), dtype=np.uint8)
        L.check(L.lib().clay_splat_tiles_download(self._h, mip, out.ctypes.data))
        return out

    def destroy(self) -> None:
        if self._owned and self._h is not None:
            L.lib().clay_splat_destroy(self._h)
        self._h = None
