"""Root (the TgpuRoot stand-in) and SdfVolume (lib/src/field/volume.ts:91-343)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .brush import brush_array

TILE = 8  # volume.ts:6


class Root:
    """One HIP device and its default stream: what `tgpu.init()` gives the reference."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        L.check(L.lib().clay_ctx_create(device, C.byref(h)))
        self._h = h
        self.device = device

    @property
    def handle(self) -> C.c_void_p:
        if self._h is None:
            raise L.ClayError(L.ERR_STATE, "Root: destroyed")
        return self._h

    def synchronize(self) -> None:
        L.check(L.lib().clay_ctx_synchronize(self.handle))

    VALU_FORMS = ("fmac", "fmac_sgpr", "fma_vop3", "fma_mix", "max", "cndmask", "rsq", "add_u32", "mul_f32")

    def valu_peak(self, form: str = "fmac", rounds: int = 16384) -> float:
        """Measured wave64 VALU instructions per second of one instruction form
        (independent register chains at full occupancy); "fmac" is the full-rate ceiling."""
        if form not in self.VALU_FORMS:
            raise ValueError(f"unknown VALU form {form!r}; one of {self.VALU_FORMS}")
        out = C.c_double()
        L.check(L.lib().clay_valu_peak(self.handle, self.VALU_FORMS.index(form), int(rounds), C.byref(out)))
        return out.value

    def atomic_peak(self, kind: str, rounds: int = 4096) -> float:
        """Measured issue rate (ops/s) of 'lds_min' (ds_min_u32, conflict-free) or
        'global_min' (global atomicMin, a wave per 256-byte span)."""
        out = C.c_double()
        k = {"lds_min": 0, "global_min": 1}[kind]
        L.check(L.lib().clay_atomic_peak(self.handle, k, int(rounds), C.byref(out)))
        return out.value

    def set_timing(self, enabled: bool) -> None:
        L.check(L.lib().clay_ctx_set_timing(self.handle, 1 if enabled else 0))

    def reset_stats(self) -> None:
        L.check(L.lib().clay_ctx_reset_stats(self.handle))

    def kernel_stats(self) -> dict[str, tuple[float, int]]:
        """name -> (total ms, launches) for every kernel timed since the last reset."""
        out = {}
        buf = C.create_string_buffer(64)
        ms = C.c_double()
        n = C.c_uint64()
        i = 0
        while L.lib().clay_ctx_kernel_stats(self.handle, i, buf, 64, C.byref(ms), C.byref(n)) == L.OK:
            out[buf.value.decode()] = (ms.value, n.value)
            i += 1
        return out

    def destroy(self) -> None:
        if self._h is not None:
            L.lib().clay_ctx_destroy(self._h)
            self._h = None


def derive(resolution: int, mip_levels: int | None = None) -> L.VolumeInfo:
    """SdfVolume's geometry and resolution check, without a device (volume.ts:139-179)."""
    info = L.VolumeInfo()
    L.check(L.lib().clay_volume_derive(int(resolution), int(mip_levels or 0), C.byref(info)))
    return info


class SdfVolume:
    """A mip-mapped band-limited distance volume.  Storage is one f16 distance plane
    and one f16 material plane per mip; `download` returns the reference's
    `rgba16float` texels (distance/band, material, 0, 1)."""

    def __init__(self, root: Root, options: dict | None = None, *, _handle=None, **kw):
        opts = dict(options or {}, **kw)
        self.root = root
        if _handle is not None:
            self._h = _handle
            self._owned = False
        else:
            d = L.VolumeDesc()
            d.resolution = int(opts.get("resolution", 128))
            d.mip_levels = int(opts.get("mipLevels", 0) or 0)
            d.world_size = float(opts.get("worldSize", 16))
            d.band = float(opts.get("band", 4))
            origin = opts.get("origin", (0, 0, 0))
            for i in range(3):
                d.origin[i] = float(origin[i])
            d.gradient_voxels = float(opts.get("gradientVoxels", 1))
            h = C.c_void_p()
            instances = int(opts.get("instances", 1))
            if instances == 1:
                L.check(L.lib().clay_volume_create(root.handle, C.byref(d), C.byref(h)))
            else:
                L.check(L.lib().clay_volume_create_instances(root.handle, C.byref(d), instances, C.byref(h)))
            self.instances = instances
            self._h = h
            self._owned = True
        self._load_info()

    def _load_info(self) -> None:
        info = L.VolumeInfo()
        L.check(L.lib().clay_volume_get_info(self._h, C.byref(info)))
        self.resolution = info.resolution
        self.mipLevels = info.mip_levels
        self.worldSize = info.world_size
        self.voxelSize = info.voxel_size
        self.band = info.band
        self.origin = tuple(info.origin)
        self.tileRes = info.tile_res[0]
        self.tileResPerMip = [info.tile_res[m] for m in range(info.mip_levels)]
        self.cellOffsets = [info.cell_offsets[m] for m in range(info.mip_levels)]
        self.totalCells = info.total_cells

    @property
    def handle(self):
        return self._h

    def setOrigin(self, origin, stream=None) -> None:
        """volume.ts:327-334."""
        L.check(L.lib().clay_volume_set_origin(self._h, L.dvec(origin), L.stream_handle(stream)))
        self.origin = tuple(float(o) for o in origin)

    def setOriginDevice(self, origin3_ptr: int, stream=None) -> None:
        """setOrigin from a device-resident float[3] (no readback)."""
        L.check(L.lib().clay_volume_set_origin_device(self._h, C.c_void_p(origin3_ptr), L.stream_handle(stream)))

    def voxelSizeAt(self, mip: int) -> float:
        out = C.c_double()
        L.check(L.lib().clay_volume_voxel_size_at(self._h, mip, C.byref(out)))
        return out.value

    def download(self, mip: int = 0, instance: int | None = None) -> np.ndarray:
        """Texels of one mip as uint16 (f16 bits), shape (r, r, r, 4) indexed [z, y, x, c];
        `instance` picks one scene's copy of a multi-instance volume."""
        r = self.resolution >> mip
        out = np.empty((r, r, r, 4), dtype=np.uint16)
        if instance is None:
            L.check(L.lib().clay_volume_download(self._h, mip, out.ctypes.data))
        else:
            L.check(L.lib().clay_volume_instance_download(self._h, int(instance), mip, out.ctypes.data))
        return out

    def replicate(self, stream=None) -> None:
        """Copies instance 0 (what bakeBrushes / upload wrote) to every instance."""
        L.check(L.lib().clay_volume_replicate(self._h, L.stream_handle(stream)))

    def upload(self, mip: int, rgba16: np.ndarray) -> None:
        r = self.resolution >> mip
        arr = np.ascontiguousarray(rgba16, dtype=np.uint16)
        if arr.size != r * r * r * 4:
            raise ValueError(f"upload: expected {r}^3 x 4 halves, got {arr.size}")
        L.check(L.lib().clay_volume_upload(self._h, mip, arr.ctypes.data))

    def packBound(self) -> int:
        """Largest stream this volume's shape can produce (every tile stored)."""
        out = C.c_uint64()
        L.check(L.lib().clay_volume_pack_bound(self._h, C.byref(out)))
        return out.value

    def pack_device(self, out=None, instance: int = 0, stream=None, sync: bool = True):
        """Packs one instance into a CUDA uint8 tensor (allocated at packBound when
        `out` is None).  Returns (tensor, length); length is None when `sync` is False."""
        import torch

        cap = self.packBound()
        if out is None:
            out = torch.empty(cap, dtype=torch.uint8, device="cuda")
        n = C.c_uint64()
        L.check(L.lib().clay_volume_pack(self._h, int(instance), C.c_void_p(out.data_ptr()), out.numel(),
                                         C.byref(n) if sync else None, L.stream_handle(stream)))
        return out, (n.value if sync else None)

    def pack(self, instance: int = 0) -> bytes:
        """The stream of one instance, as host bytes (a file's contents)."""
        t, n = self.pack_device(instance=instance)
        return t[:n].cpu().numpy().tobytes()

    def unpack_device(self, src, length: int, instance: int = 0, stream=None, header=None) -> None:
        """From a CUDA uint8 tensor.  With `header` (a WireHeader already read) nothing
        synchronises and a bad table word is reported later by `wireStatus`."""
        L.check(L.lib().clay_volume_unpack(self._h, int(instance), C.c_void_p(src.data_ptr()), int(length),
                                           None if header is None else C.byref(header), L.stream_handle(stream)))

    def wireStatus(self) -> int:
        bad = C.c_uint32()
        L.check(L.lib().clay_volume_wire_status(self._h, C.byref(bad)))
        return bad.value

    def unpack(self, data: bytes, instance: int = 0) -> None:
        """Replaces one instance's texels, every mip, with a stream from host bytes."""
        import torch

        src = torch.frombuffer(bytearray(data), dtype=torch.uint8).cuda()
        self.unpack_device(src, len(data), instance)

    def bakeBrushes(self, brushes, stream=None) -> None:
        """SdfBuilder.setBrushes + rebuild (builder.ts:145-172)."""
        arr = brush_array(brushes)
        L.check(L.lib().clay_volume_bake_brushes(self._h, arr, len(arr), L.stream_handle(stream)))

    def destroy(self) -> None:
        if self._owned and self._h is not None:
            L.lib().clay_volume_destroy(self._h)
        self._h = None
