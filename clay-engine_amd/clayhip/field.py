"""TracedField descriptors (lib/src/trace/field.ts).

The reference composes shader closures; the C ABI composes descriptors that the
integrate kernel interprets: an analytic brush fold, a baked volume, or a union
of up to four of those (unionField folded left).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .brush import brush_array
from .volume import Root, SdfVolume


class TracedField:
    def __init__(self, root: Root, handle, owned: bool = True, keep=()):
        self.root = root
        self._h = handle
        self._owned = owned
        self._keep = tuple(keep)  # operands that must outlive this field

    @property
    def handle(self):
        return self._h

    @property
    def maxMip(self) -> int:
        out = C.c_int32()
        L.check(L.lib().clay_field_max_mip(self._h, C.byref(out)))
        return out.value

    def evaluate(self, points) -> tuple[np.ndarray, np.ndarray]:
        """`(sample(p, 0).x, normal(p))` at host points (n, 3), computed on the GPU."""
        pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
        n = pts.shape[0]
        dist = np.empty(n, dtype=np.float32)
        nrm = np.empty((n, 3), dtype=np.float32)
        L.check(L.lib().clay_field_eval(self._h, pts.ctypes.data, n, dist.ctypes.data, nrm.ctypes.data))
        return dist, nrm

    def _query(self, what: int, points, mip: int) -> np.ndarray:
        pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
        out = np.empty((pts.shape[0], 2), dtype=np.float32)
        L.check(L.lib().clay_field_query(self._h, what, int(mip), pts.ctypes.data, pts.shape[0], out.ctypes.data))
        return out

    def sample(self, points, mip: int = 0) -> np.ndarray:
        """TracedField.sample(p, mip) (field.ts:35-41): (n, 2) world distance and
        band-normalised value, computed on the GPU."""
        return self._query(L.QUERY_SAMPLE, points, mip)

    def field(self, points, mip: int = 0) -> np.ndarray:
        """TracedField.field(p, mip) (field.ts:42-43): (n, 2) world distance and material."""
        return self._query(L.QUERY_FIELD, points, mip)

    def entry(self, ro, rd, tMin=0.0, radius=0.0, aperture=0.0) -> np.ndarray:
        """TracedField.entry (field.ts:55-74) per ray: the first t at which a swept
        shape could touch the field, or 1e9 (NEVER)."""
        ro = np.asarray(ro, dtype=np.float32).reshape(-1, 3)
        rd = np.asarray(rd, dtype=np.float32).reshape(-1, 3)
        n = ro.shape[0]
        rays = np.empty((n, 9), dtype=np.float32)
        rays[:, 0:3], rays[:, 3:6] = ro, rd
        rays[:, 6], rays[:, 7], rays[:, 8] = tMin, radius, aperture
        out = np.empty(n, dtype=np.float32)
        L.check(L.lib().clay_field_entry(self._h, rays.ctypes.data, n, out.ctypes.data))
        return out

    def sweep(self, ro, rd, tMin=0.0, tMax=1e9, radius=0.0, aperture=0.0, options: dict | None = None) -> dict:
        """makeTracer(field, options).sweep (march.ts:97-218) per ray, on the GPU: ray
        (radius and aperture 0), sphere sweep (radius) or cone (aperture).  The
        per-ray arguments broadcast; `options` takes maxSteps, minStep, hitEps and
        minMip.  Returns the SweepHit fields as arrays: t, hit, steps, dist."""
        ro = np.asarray(ro, dtype=np.float32).reshape(-1, 3)
        rd = np.asarray(rd, dtype=np.float32).reshape(-1, 3)
        n = max(ro.shape[0], rd.shape[0])
        rays = np.empty((n, 10), dtype=np.float32)
        rays[:, 0:3], rays[:, 3:6] = ro, rd
        rays[:, 6], rays[:, 7], rays[:, 8], rays[:, 9] = tMin, tMax, radius, aperture
        hits = np.empty(n, dtype=[("t", np.float32), ("hit", np.uint32), ("steps", np.uint32), ("dist", np.float32)])
        L.check(L.lib().clay_field_sweep(self._h, C.byref(tracer_opts(options)), rays.ctypes.data, n,
                                         hits.ctypes.data))
        return {k: hits[k].copy() for k in ("t", "hit", "steps", "dist")}

    def sweep_device(self, rays_ptr: int, n: int, hits_ptr: int, stream: int = 0, options: dict | None = None) -> None:
        """The sweep over device buffers (10 floats per ray in, 16-byte SweepHit out),
        enqueued on `stream` (0: the context's own)."""
        L.check(L.lib().clay_field_sweep_device(self._h, C.byref(tracer_opts(options)), C.c_void_p(rays_ptr), int(n),
                                                C.c_void_p(hits_ptr), C.c_void_p(stream or None)))

    def bandWorld(self, mip: int = 0) -> float:
        b = C.c_float()
        L.check(L.lib().clay_field_scale(self._h, int(mip), C.byref(b), None))
        return b.value

    def voxelWorld(self, mip: int = 0) -> float:
        v = C.c_float()
        L.check(L.lib().clay_field_scale(self._h, int(mip), None, C.byref(v)))
        return v.value

    def destroy(self) -> None:
        if self._owned and self._h is not None:
            L.lib().clay_field_destroy(self._h)
        self._h = None


def tracer_opts(options: dict | None = None) -> L.TracerOpts:
    """makeTracer's options (march.ts:22-40) with the reference's defaults."""
    o = L.TracerOpts()
    L.lib().clay_tracer_opts_default(C.byref(o))
    for key, field in (("maxSteps", "max_steps"), ("minStep", "min_step"), ("hitEps", "hit_eps"), ("minMip", "min_mip")):
        if options and options.get(key) is not None:
            setattr(o, field, options[key])
    return o


def analyticField(root: Root, brushes, options: dict | None = None) -> TracedField:
    """analyticField (field.ts:196-236) over the fold of `brushes` (BrushDesc dicts
    or Brush records) from (1e9, 0)."""
    opts = options or {}
    arr = brush_array(brushes)
    o = L.AnalyticOpts()
    o.band = float(opts.get("band", 64))
    o.epsilon = float(opts.get("epsilon", 1e-3))
    o.material = float(opts.get("material", 0))
    h = C.c_void_p()
    L.check(L.lib().clay_field_analytic(root.handle, arr, len(arr), C.byref(o), C.byref(h)))
    return TracedField(root, h)


def volumeField(volume: SdfVolume) -> TracedField:
    """volumeField (field.ts:102-180)."""
    h = C.c_void_p()
    L.check(L.lib().clay_field_volume(volume.root.handle, volume.handle, C.byref(h)))
    return TracedField(volume.root, h, keep=(volume,))


def unionField(a: TracedField, b: TracedField) -> TracedField:
    """unionField (field.ts:246-315)."""
    h = C.c_void_p()
    L.check(L.lib().clay_field_union(a.handle, b.handle, C.byref(h)))
    return TracedField(a.root, h, keep=(a, b))


def _device_floats(value, n: int):
    """A CUDA float32 tensor of n values: `value` itself when it is one (the caller
    may then change it in place, as the reference's GPU-side closure reads a
    uniform), else a new one holding it."""
    import torch

    if isinstance(value, torch.Tensor):
        if not (value.is_cuda and value.dtype == torch.float32 and value.numel() == n and value.is_contiguous()):
            raise ValueError(f"expected a contiguous CUDA float32 tensor of {n} values")
        return value
    return torch.tensor(np.asarray(value, dtype=np.float32).reshape(n), device="cuda")


def offsetField(field: TracedField, offset) -> TracedField:
    """offsetField (field.ts:388-429): `field` read at p - offset.  `offset` is a
    CUDA tensor of 3 floats read on the device at every evaluation, or 3 numbers."""
    dev = _device_floats(offset, 3)
    h = C.c_void_p()
    L.check(L.lib().clay_field_offset(field.handle, C.c_void_p(dev.data_ptr()), C.byref(h)))
    out = TracedField(field.root, h, keep=(field, dev))
    out.offset = dev
    return out


def lerpField(a: TracedField, b: TracedField, t) -> TracedField:
    """lerpField (field.ts:316-386) of a morph pair of volume fields, mixed by
    clamp(t, 0, 1); `t` is a CUDA tensor of one float read on the device, or a number."""
    dev = _device_floats(t, 1)
    h = C.c_void_p()
    L.check(L.lib().clay_field_lerp(a.handle, b.handle, C.c_void_p(dev.data_ptr()), C.byref(h)))
    out = TracedField(a.root, h, keep=(a, b, dev))
    out.t = dev
    return out
