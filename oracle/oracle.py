"""ctypes wrapper of the CPU oracle (oracle/libclay_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, bench.py's cpu_baseline leg and
__graft_entry__.smoke() as the checker.  The product (clay-engine_amd/) never
imports it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "libclay_oracle.so"
INVALID_KEY = 0xFFFFFFFF
CONTACTS = 32


def build() -> Path:
    res = subprocess.run(["make", "-s", "-C", str(HERE)], capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{res.stdout}\n{res.stderr}")
    return LIB


class OrcBrush(C.Structure):
    _fields_ = [
        ("pos", C.c_float * 3), ("scale", C.c_float), ("rot", C.c_float * 4), ("size", C.c_float * 3),
        ("radius", C.c_float), ("material", C.c_float), ("blend", C.c_float), ("bound", C.c_float),
        ("mask", C.c_float), ("kind", C.c_uint32), ("op", C.c_uint32), ("slot", C.c_uint32),
        ("pad", C.c_uint32),
    ]


class OrcFluidOpts(C.Structure):
    _fields_ = [
        ("spacing", C.c_double), ("smoothing", C.c_double), ("density", C.c_double), ("radius", C.c_double),
        ("bucket", C.c_int32), ("substeps", C.c_int32), ("origin", C.c_double * 3), ("world_size", C.c_double),
        ("gravity", C.c_double * 3), ("stiffness", C.c_double), ("viscosity", C.c_double),
        ("damping", C.c_double), ("max_speed", C.c_double),
    ]


class OrcFluidK(C.Structure):
    _fields_ = [
        ("h", C.c_double), ("mass", C.c_double), ("poly6", C.c_double), ("spiky_grad", C.c_double),
        ("rest_density", C.c_double), ("dt", C.c_double), ("radius", C.c_double), ("contact_depth", C.c_double),
        ("grid_origin", C.c_double * 3), ("grid_res", C.c_int32), ("bucket", C.c_int32), ("substeps", C.c_int32),
        ("pad0", C.c_int32), ("wall_lo", C.c_double * 3), ("wall_hi", C.c_double * 3), ("gravity", C.c_double * 3),
        ("stiffness", C.c_double), ("viscosity", C.c_double), ("damping", C.c_double), ("max_speed", C.c_double),
        ("hh", C.c_double), ("mass_poly6", C.c_double), ("h6", C.c_double), ("inv_h", C.c_double),
        ("rest_floor", C.c_double),
    ]


class OrcLeaf(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("nbrush", C.c_int32), ("brushes", C.c_void_p), ("eps", C.c_double),
        ("dist16", C.c_void_p), ("res", C.c_int32), ("pad1", C.c_int32), ("origin", C.c_double * 3),
        ("world_size", C.c_double), ("band", C.c_double), ("grad_voxels", C.c_double),
        ("mat16", C.c_void_p), ("material", C.c_double),
        ("mips", C.c_int32), ("pad2", C.c_int32), ("dist_m", C.c_void_p * 8), ("mat_m", C.c_void_p * 8),
        ("analytic_band", C.c_double),
        ("has_offset", C.c_int32), ("pad3", C.c_int32), ("offset", C.c_double * 3),
        ("lerp_b", C.c_void_p), ("lerp_t", C.c_double),
    ]


class OrcCollider(C.Structure):
    _fields_ = [("nleaf", C.c_int32), ("pad", C.c_int32), ("leaf", OrcLeaf * 4)]


class OrcTracerOpts(C.Structure):
    _fields_ = [("max_steps", C.c_int32), ("min_mip", C.c_int32), ("min_step", C.c_double), ("hit_eps", C.c_double)]


class OrcSplatOpts(C.Structure):
    _fields_ = [
        ("res", C.c_int32), ("mips", C.c_int32), ("world_size", C.c_double), ("origin", C.c_double * 3),
        ("band", C.c_double), ("radius", C.c_double), ("material", C.c_double),
    ]


class OrcState(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("px", "py", "pz", "vx", "vy", "vz", "rho", "pres", "nvx", "nvy", "nvz")]


class OrcContacts(C.Structure):
    _fields_ = [("count", C.c_uint32), ("pad", C.c_uint32), ("points", (C.c_double * 4) * CONTACTS)]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        L.orc_hash_u32.restype = C.c_uint32
        L.orc_hash_u32.argtypes = [C.c_uint32]
        L.orc_hash1.restype = C.c_float
        L.orc_hash1.argtypes = [C.c_uint32]
        L.orc_f32_to_f16.restype = C.c_uint16
        L.orc_f32_to_f16.argtypes = [C.c_float]
        L.orc_f16_to_f32.restype = C.c_float
        L.orc_f16_to_f32.argtypes = [C.c_uint16]
        L.orc_enc_f32.restype = C.c_uint32
        L.orc_enc_f32.argtypes = [C.c_float]
        L.orc_dec_f32.restype = C.c_float
        L.orc_dec_f32.argtypes = [C.c_uint32]
        L.orc_pack_tile.restype = C.c_uint32
        L.orc_pack_tile.argtypes = [C.c_uint32] * 3
        L.orc_volume_derive.restype = C.c_int
        L.orc_volume_derive.argtypes = [C.c_int, C.c_int] + [C.c_void_p] * 5 + [C.c_int]
        L.orc_get_threads.restype = C.c_int
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_unpack_tile.argtypes = [C.c_uint32, C.c_void_p]
        L.orc_fluid_derive.argtypes = [C.POINTER(OrcFluidOpts), C.POINTER(OrcFluidK)]
        L.orc_stable_sort_keys.argtypes = [C.c_int, C.c_void_p, C.c_void_p]
        P, I, K = C.c_void_p, C.c_int, C.POINTER(OrcFluidK)
        COL, ST, CT = C.POINTER(OrcCollider), C.POINTER(OrcState), C.POINTER(OrcContacts)
        SO = C.POINTER(OrcSplatOpts)
        for s in ("f32", "f64"):
            getattr(L, f"orc_keys_{s}").argtypes = [K, I, P, P, P, P]
            getattr(L, f"orc_substep_{s}").argtypes = [K, COL, I, ST, CT, P, P]
            getattr(L, f"orc_neighbors_{s}").argtypes = [K, I, P, P, P, I, P, I]
            getattr(L, f"orc_neighbors_{s}").restype = C.c_int
            getattr(L, f"orc_emit_{s}").argtypes = [I, I, I, C.c_uint32, P, C.c_double, P, ST]
            getattr(L, f"orc_density_{s}").argtypes = [K, I, ST]
            getattr(L, f"orc_forces_{s}").argtypes = [K, I, ST]
            getattr(L, f"orc_integrate_{s}").argtypes = [K, COL, I, ST, CT]
            getattr(L, f"orc_collider_eval_{s}").argtypes = [COL, I, P, P, P]
            getattr(L, f"orc_splat_bake_{s}").argtypes = [SO, I, P, P, P, P, P, P, I, I, P]
            getattr(L, f"orc_splat_resolve_{s}").argtypes = [SO, P, P]
            getattr(L, f"orc_bake_brushes_{s}").argtypes = [I, I, C.c_double, P, C.c_double, P, I, P]
            D = C.c_double
            getattr(L, f"orc_edit_bin_{s}").argtypes = [I, D, P, D, I, P, I, P, P]
            getattr(L, f"orc_edit_apply_{s}").argtypes = [I, D, P, D, I, P, P, P, P, P]
            getattr(L, f"orc_mip_refresh_{s}").argtypes = [I, I, I, D, D, P, P, P]
            getattr(L, f"orc_edit_flush_{s}").argtypes = [I, I, D, P, D, D, I, P, I, P, P, P]
            getattr(L, f"orc_body_centroids_{s}").argtypes = [I, P, P, P, P, P, P, P, P]
            getattr(L, f"orc_pbd_step_{s}").argtypes = [C.POINTER(OrcPbdParams), COL, I, I, I, P, P, P, P, P, I, I]
            getattr(L, f"orc_sin_cos_{s}").argtypes = [C.c_float if s == "f32" else C.c_double, P, P]
            getattr(L, f"orc_field_query_{s}").argtypes = [COL, I, I, I, P, P]
            getattr(L, f"orc_field_entry_{s}").argtypes = [COL, I, P, P]
            getattr(L, f"orc_field_sweep_{s}").argtypes = [COL, C.POINTER(OrcTracerOpts), I, P, P, P]
            getattr(L, f"orc_extract_{s}").argtypes = [C.POINTER(OrcExtractParams), COL, I, P, P, P, P, P, P, P,
                                                       P, P, P]
        L.orc_tile_dilate.argtypes = [I, I, I, P]
        L.orc_contacts_to_edits.argtypes = [C.c_uint32, P, C.c_double, C.c_double, C.c_double, I, P]
        L.orc_contacts_to_edits.restype = C.c_int
        _lib = L
    return _lib


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def threads() -> int:
    return lib().orc_get_threads()


# ---- scalar helpers ------------------------------------------------------------
def hash_u32(x: int) -> int:
    return lib().orc_hash_u32(x)


def hash1(x: int) -> float:
    return lib().orc_hash1(x)


def enc(v: float) -> int:
    return lib().orc_enc_f32(v)


def dec(e: int) -> float:
    return lib().orc_dec_f32(e)


def f32_to_f16(v: float) -> int:
    return lib().orc_f32_to_f16(v)


def f16_to_f32(h: int) -> float:
    return lib().orc_f16_to_f32(h)


def pack_tile(x: int, y: int, z: int) -> int:
    return lib().orc_pack_tile(x, y, z)


def unpack_tile(v: int) -> tuple[int, int, int]:
    out = (C.c_uint32 * 3)()
    lib().orc_unpack_tile(C.c_uint32(v), out)
    return tuple(out)


def volume_derive(resolution: int, mips: int = 0):
    """-> (mips, tile_res list, cell offsets, total cells) or raises ValueError(message)."""
    m = C.c_int()
    tr = (C.c_int * 32)()
    co = (C.c_int * 32)()
    tot = C.c_int()
    err = C.create_string_buffer(256)
    rc = lib().orc_volume_derive(resolution, mips, C.byref(m), tr, co, C.byref(tot), err, 256)
    if rc != 0:
        raise ValueError(err.value.decode())
    return m.value, list(tr[: m.value]), list(co[: m.value]), tot.value


def stable_sort(keys: np.ndarray) -> np.ndarray:
    k = np.ascontiguousarray(keys, dtype=np.uint32)
    out = np.empty(k.size, dtype=np.uint32)
    lib().orc_stable_sort_keys(k.size, k.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p))
    return out


# ---- fluid ---------------------------------------------------------------------
DEFAULTS = {
    "spacing": 0.45, "smoothing": -1.0, "density": 1.0, "radius": -1.0, "bucket": 64, "substeps": 3,
    "origin": (-12.0, -4.0, -12.0), "worldSize": 24.0, "gravity": (0.0, -9.81, 0.0), "stiffness": 220.0,
    "viscosity": 0.4, "damping": 0.995, "maxSpeed": 15.0,
}


def fluid_consts(options: dict) -> OrcFluidK:
    o = dict(DEFAULTS)
    o.update({k: v for k, v in options.items() if k in DEFAULTS})
    op = OrcFluidOpts()
    op.spacing = o["spacing"]
    op.smoothing = o["smoothing"]
    op.density = o["density"]
    op.radius = o["radius"]
    op.bucket = int(o["bucket"])
    op.substeps = int(o["substeps"])
    for i in range(3):
        op.origin[i] = o["origin"][i]
        op.gravity[i] = o["gravity"][i]
    op.world_size = o["worldSize"]
    op.stiffness = o["stiffness"]
    op.viscosity = o["viscosity"]
    op.damping = o["damping"]
    op.max_speed = o["maxSpeed"]
    k = OrcFluidK()
    lib().orc_fluid_derive(C.byref(op), C.byref(k))
    return k


class Collider:
    """Union (left fold) of analytic brush folds and mip-0 volume planes."""

    def __init__(self):
        self.c = OrcCollider()
        self._keep = []

    def add_analytic(self, brush_records, eps: float = 1e-3, material: float = 0.0, band: float = 64.0) -> "Collider":
        arr = (OrcBrush * len(brush_records))()
        for i, b in enumerate(brush_records):
            C.memmove(C.byref(arr[i]), C.byref(b), C.sizeof(OrcBrush))
        self._keep.append(arr)
        leaf = self.c.leaf[self.c.nleaf]
        leaf.kind = 0
        leaf.nbrush = len(brush_records)
        leaf.brushes = C.cast(arr, C.c_void_p)
        leaf.eps = eps
        leaf.material = material
        leaf.analytic_band = band
        leaf.mips = 1
        self.c.nleaf += 1
        return self

    def add_volume(self, dist16: np.ndarray, res: int, origin, world_size: float, band: float,
                   grad_voxels: float = 1.0, mat16: np.ndarray | None = None, levels=None) -> "Collider":
        """`mat16`: the mip-0 material channel (f16 bits); None reads material 0.
        `levels`: every mip as (dist16, mat16) pairs, for the reads at any level."""
        d = np.ascontiguousarray(dist16, dtype=np.uint16).reshape(-1)
        assert d.size == res ** 3
        self._keep.append(d)
        leaf = self.c.leaf[self.c.nleaf]
        leaf.kind = 1
        leaf.dist16 = d.ctypes.data
        leaf.res = res
        for i in range(3):
            leaf.origin[i] = origin[i]
        leaf.world_size = world_size
        leaf.band = band
        leaf.grad_voxels = grad_voxels
        if mat16 is not None:
            m = np.ascontiguousarray(mat16, dtype=np.uint16).reshape(-1)
            assert m.size == res ** 3
            self._keep.append(m)
            leaf.mat16 = m.ctypes.data
        leaf.mips = 1
        if levels is not None:
            leaf.mips = len(levels)
            for lv, (dl, ml) in enumerate(levels):
                dl = np.ascontiguousarray(dl, dtype=np.uint16).reshape(-1)
                ml = np.ascontiguousarray(ml, dtype=np.uint16).reshape(-1)
                assert dl.size == (res >> lv) ** 3 and ml.size == dl.size
                self._keep += [dl, ml]
                leaf.dist_m[lv] = dl.ctypes.data
                leaf.mat_m[lv] = ml.ctypes.data
        self.c.nleaf += 1
        return self

    def offset(self, o) -> "Collider":
        """offsetField over every leaf so far (field.ts:388-429): each reads p - o."""
        for i in range(self.c.nleaf):
            leaf = self.c.leaf[i]
            assert not leaf.has_offset, "one offset per leaf"
            leaf.has_offset = 1
            for a in range(3):
                leaf.offset[a] = float(o[a])
        return self

    def add_lerp(self, a: "Collider", b: "Collider", t: float) -> "Collider":
        """lerpField(volumeField(A), volumeField(B), t) (field.ts:316-386) as one leaf,
        from two one-volume colliders."""
        assert a.c.nleaf == 1 and b.c.nleaf == 1 and a.c.leaf[0].kind == 1 and b.c.leaf[0].kind == 1
        second = OrcLeaf()
        C.memmove(C.byref(second), C.byref(b.c.leaf[0]), C.sizeof(OrcLeaf))
        self._keep += [a, b, second]
        leaf = self.c.leaf[self.c.nleaf]
        C.memmove(C.byref(leaf), C.byref(a.c.leaf[0]), C.sizeof(OrcLeaf))
        leaf.kind = 2
        leaf.lerp_b = C.cast(C.pointer(second), C.c_void_p)
        leaf.lerp_t = float(t)
        self.c.nleaf += 1
        return self

    def query(self, points: np.ndarray, mip: int = 0, what: str = "sample", precision: str = "f32"):
        """TracedField.sample / field at `mip`: (n, 2)."""
        dt = np.float32 if precision == "f32" else np.float64
        p = np.ascontiguousarray(points, dtype=dt).reshape(-1, 3)
        out = np.empty((p.shape[0], 2), dtype=dt)
        getattr(lib(), f"orc_field_query_{precision}")(C.byref(self.c), 1 if what == "field" else 0, mip,
                                                       p.shape[0], p.ctypes.data, out.ctypes.data)
        return out

    def entry(self, rays9: np.ndarray, precision: str = "f32") -> np.ndarray:
        dt = np.float32 if precision == "f32" else np.float64
        r = np.ascontiguousarray(rays9, dtype=dt).reshape(-1, 9)
        out = np.empty(r.shape[0], dtype=dt)
        getattr(lib(), f"orc_field_entry_{precision}")(C.byref(self.c), r.shape[0], r.ctypes.data, out.ctypes.data)
        return out

    def sweep(self, rays10: np.ndarray, max_steps: int = 192, min_step: float = 0.25, hit_eps: float = 0.5,
              min_mip: int = 0, precision: str = "f32") -> dict:
        """makeTracer(field, options).sweep per ray (march.ts:97-218); rays10 rows are
        ro xyz, rd xyz, tMin, tMax, radius, aperture.  Returns t, hit, steps, dist."""
        dt = np.float32 if precision == "f32" else np.float64
        r = np.ascontiguousarray(rays10, dtype=dt).reshape(-1, 10)
        n = r.shape[0]
        td = np.empty((n, 2), dtype=dt)
        hs = np.empty((n, 2), dtype=np.uint32)
        o = OrcTracerOpts(max_steps=max_steps, min_mip=min_mip, min_step=min_step, hit_eps=hit_eps)
        getattr(lib(), f"orc_field_sweep_{precision}")(C.byref(self.c), C.byref(o), n, r.ctypes.data, td.ctypes.data,
                                                       hs.ctypes.data)
        return {"t": td[:, 0].copy(), "hit": hs[:, 0].copy(), "steps": hs[:, 1].copy(), "dist": td[:, 1].copy()}

    def eval(self, points: np.ndarray, precision: str = "f32"):
        dt = np.float32 if precision == "f32" else np.float64
        p = np.ascontiguousarray(points, dtype=dt).reshape(-1, 3)
        n = p.shape[0]
        dist = np.empty(n, dtype=dt)
        nrm = np.empty((n, 3), dtype=dt)
        getattr(lib(), f"orc_collider_eval_{precision}")(C.byref(self.c), n, p.ctypes.data_as(C.c_void_p),
                                                         dist.ctypes.data_as(C.c_void_p),
                                                         nrm.ctypes.data_as(C.c_void_p))
        return dist, nrm


class Fluid:
    """Oracle state of one FluidSim: SoA arrays of `capacity` entries in f32
    ("WGSL-f32") or f64, the live prefix `n`, and the emission ring bookkeeping of
    FluidSim.spawn/step (fluid.ts:511-549)."""

    FIELDS = ("px", "py", "pz", "vx", "vy", "vz", "rho", "pres", "nvx", "nvy", "nvz")

    def __init__(self, options: dict, pos: np.ndarray | None = None, vel: np.ndarray | None = None,
                 collider: Collider | None = None, precision: str = "f32", capacity: int | None = None):
        self.precision = precision
        self.dt = np.float32 if precision == "f32" else np.float64
        self.options = dict(options)
        self.k = fluid_consts(options)
        self.collider = collider
        self.n = 0 if pos is None else pos.shape[0]
        self.capacity = capacity if capacity is not None else max(self.n, 1)
        self.a = {f: np.zeros(self.capacity, dtype=self.dt) for f in self.FIELDS}
        for i, ax in enumerate("xyz"):
            if pos is not None:
                self.a["p" + ax][: self.n] = pos[:, i]
            if vel is not None:
                self.a["v" + ax][: self.n] = vel[:, i]
        self.contacts = OrcContacts()
        self.keys = np.zeros(self.capacity, dtype=np.uint32)
        self.perm = np.zeros(self.capacity, dtype=np.uint32)
        self.cursor = self.n % self.capacity
        self.pending = 0
        self.seed = 1
        self.emit_args = None

    def _state(self) -> OrcState:
        st = OrcState()
        for f in self.FIELDS:
            setattr(st, f, self.a[f].ctypes.data)
        return st

    def substep(self) -> None:
        st = self._state()
        col = C.byref(self.collider.c) if self.collider else None
        getattr(lib(), f"orc_substep_{self.precision}")(
            C.byref(self.k), col, self.n, C.byref(st), C.byref(self.contacts),
            self.keys.ctypes.data_as(C.c_void_p), self.perm.ctypes.data_as(C.c_void_p))

    def density(self) -> None:
        """The density pass alone (fluid.ts:333-368) over the current positions."""
        st = self._state()
        getattr(lib(), f"orc_density_{self.precision}")(C.byref(self.k), self.n, C.byref(st))

    def forces(self) -> None:
        """The force pass alone (fluid.ts:370-439): writes nvel from the current
        positions, velocities, densities and pressures."""
        st = self._state()
        getattr(lib(), f"orc_forces_{self.precision}")(C.byref(self.k), self.n, C.byref(st))

    def integrate(self) -> None:
        """The integrate pass alone (fluid.ts:444-491): advances by nvel, collides,
        records contacts and clamps to the walls."""
        st = self._state()
        col = C.byref(self.collider.c) if self.collider else None
        self.contacts = OrcContacts()
        getattr(lib(), f"orc_integrate_{self.precision}")(C.byref(self.k), col, self.n, C.byref(st),
                                                          C.byref(self.contacts))

    def set(self, what: str, values: np.ndarray) -> None:
        """Overwrites the first n entries of a field: 'pos', 'vel', 'nvel' (n, 3) or
        'density', 'pressure' (n,)."""
        n = self.n
        if what in ("pos", "vel", "nvel"):
            pre = {"pos": "p", "vel": "v", "nvel": "nv"}[what]
            for i, ax in enumerate("xyz"):
                self.a[pre + ax][:n] = values[:, i]
        else:
            self.a[{"density": "rho", "pressure": "pres"}[what]][:n] = values

    def spawn(self, count: int, pos, vel, spread: float | None = None) -> None:
        """FluidSim.spawn (fluid.ts:511-525)."""
        spacing = self.options.get("spacing", DEFAULTS["spacing"])
        if spread is None:
            spread = spacing * max(count, 1) ** (1.0 / 3.0) * 0.62
        self.pending = min(count, self.capacity)
        self.emit_args = (tuple(map(float, pos)), float(spread), tuple(map(float, vel)), self.seed)
        self.seed += 1

    def emit_pending(self) -> None:
        """The emission half of FluidSim.step (fluid.ts:537-549)."""
        count, self.pending = self.pending, 0
        if count <= 0:
            return
        base = self.cursor
        self.cursor = (base + count) % self.capacity
        self.n = min(self.capacity, max(self.n, base + count))
        pos, spread, vel, seed = self.emit_args
        st = self._state()
        getattr(lib(), f"orc_emit_{self.precision}")(
            self.capacity, base, count, seed, (C.c_double * 3)(*pos), C.c_double(spread),
            (C.c_double * 3)(*vel), C.byref(st))

    def get(self, what: str) -> np.ndarray:
        n = self.n
        if what in ("pos", "vel", "nvel"):
            pre = {"pos": "p", "vel": "v", "nvel": "nv"}[what]
            return np.stack([self.a[pre + ax][:n] for ax in "xyz"], axis=-1)
        return self.a[{"density": "rho", "pressure": "pres"}[what]][:n].copy()

    def read_contacts(self):
        n = min(self.contacts.count, CONTACTS)
        return self.contacts.count, [tuple(self.contacts.points[i]) for i in range(n)]


# ---- splat ---------------------------------------------------------------------
def splat_opts(res: int, mips: int, world_size: float, origin, band: float, radius: float,
               material: float = 0.0) -> OrcSplatOpts:
    o = OrcSplatOpts()
    o.res, o.mips, o.world_size = res, mips, world_size
    for i in range(3):
        o.origin[i] = origin[i]
    o.band, o.radius, o.material = band, radius, material
    return o


def grid_sizes(res: int, mips: int) -> list[int]:
    return [(res >> m) ** 3 for m in range(mips)]


def splat_bake(o: OrcSplatOpts, pos: np.ndarray, live_first: int = 0, live_count: int | None = None,
               normals: np.ndarray | None = None, precision: str = "f32") -> list[np.ndarray]:
    """Per-mip u32 code grids, each shaped (r, r, r) indexed [z, y, x]."""
    dt = np.float32 if precision == "f32" else np.float64
    p = np.ascontiguousarray(np.asarray(pos, dtype=dt).T)
    n = p.shape[1]
    if live_count is None:
        live_count = n
    total = sum(grid_sizes(o.res, o.mips))
    grid = np.empty(total, dtype=np.uint32)
    nv = [None, None, None]
    if normals is not None:
        nn = np.ascontiguousarray(np.asarray(normals, dtype=dt).T)
        nv = [nn[i].ctypes.data for i in range(3)]
    getattr(lib(), f"orc_splat_bake_{precision}")(
        C.byref(o), n, p[0].ctypes.data, p[1].ctypes.data, p[2].ctypes.data,
        C.c_void_p(nv[0]), C.c_void_p(nv[1]), C.c_void_p(nv[2]), live_first, live_count, grid.ctypes.data)
    out, off = [], 0
    for m in range(o.mips):
        r = o.res >> m
        out.append(grid[off:off + r ** 3].reshape(r, r, r))
        off += r ** 3
    return out


def splat_resolve(o: OrcSplatOpts, grids: list[np.ndarray], precision: str = "f32") -> list[np.ndarray]:
    flat = np.concatenate([g.reshape(-1) for g in grids]).astype(np.uint32)
    tex = np.empty(flat.size * 4, dtype=np.uint16)
    getattr(lib(), f"orc_splat_resolve_{precision}")(C.byref(o), flat.ctypes.data, tex.ctypes.data)
    out, off = [], 0
    for m in range(o.mips):
        r = o.res >> m
        out.append(tex[4 * off:4 * (off + r ** 3)].reshape(r, r, r, 4))
        off += r ** 3
    return out


def bake_brushes(res: int, mips: int, world_size: float, origin, band: float, brush_records,
                 precision: str = "f32") -> list[np.ndarray]:
    arr = (OrcBrush * max(1, len(brush_records)))()
    for i, b in enumerate(brush_records):
        C.memmove(C.byref(arr[i]), C.byref(b), C.sizeof(OrcBrush))
    total = sum(grid_sizes(res, mips))
    tex = np.empty(total * 4, dtype=np.uint16)
    org = (C.c_double * 3)(*origin)
    getattr(lib(), f"orc_bake_brushes_{precision}")(res, mips, C.c_double(world_size), org, C.c_double(band),
                                                    arr, len(brush_records), tex.ctypes.data)
    out, off = [], 0
    for m in range(mips):
        r = res >> m
        out.append(tex[4 * off:4 * (off + r ** 3)].reshape(r, r, r, 4))
        off += r ** 3
    return out


def neighbors(options: dict, pos: np.ndarray, i: int, precision: str = "f64") -> list[int]:
    """Ids the oracle's 27-cell scan finds within h of particle i."""
    dt = np.float32 if precision == "f32" else np.float64
    p = np.ascontiguousarray(np.asarray(pos, dtype=dt).T)
    k = fluid_consts(options)
    out = np.empty(p.shape[1], dtype=np.uint32)
    n = getattr(lib(), f"orc_neighbors_{precision}")(C.byref(k), p.shape[1], p[0].ctypes.data, p[1].ctypes.data,
                                                      p[2].ctypes.data, i, out.ctypes.data, out.size)
    return sorted(int(v) for v in out[:n])


def keys(options: dict, pos: np.ndarray) -> np.ndarray:
    k = fluid_consts(options)
    p = np.ascontiguousarray(np.asarray(pos, dtype=np.float32).T)
    out = np.empty(p.shape[1], dtype=np.uint32)
    lib().orc_keys_f32(C.byref(k), p.shape[1], p[0].ctypes.data, p[1].ctypes.data, p[2].ctypes.data,
                       out.ctypes.data)
    return out


# ---- SdfEditor / TileGrid / MipRefiner / body centroids ----------------------
def _brush_array(records):
    arr = (OrcBrush * max(1, len(records)))()
    for i, b in enumerate(records):
        C.memmove(C.byref(arr[i]), C.byref(b), C.sizeof(OrcBrush))
    return arr


class Planes:
    """A volume's distance and material planes (f16 bits, mips concatenated, x
    fastest), as the HIP library stores them."""

    def __init__(self, res: int, mips: int, texels: list[np.ndarray] | None = None):
        self.res, self.mips = res, mips
        total = sum(grid_sizes(res, mips))
        self.dist = np.empty(total, dtype=np.uint16)
        self.mat = np.empty(total, dtype=np.uint16)
        if texels is not None:
            off = 0
            for m in range(mips):
                t = np.asarray(texels[m]).reshape(-1, 4)
                self.dist[off:off + t.shape[0]] = t[:, 0]
                self.mat[off:off + t.shape[0]] = t[:, 1]
                off += t.shape[0]

    def level(self, m: int, plane: str = "dist") -> np.ndarray:
        off = sum(grid_sizes(self.res, m))
        r = self.res >> m
        return getattr(self, plane)[off:off + r ** 3].reshape(r, r, r)

    def texels(self, m: int) -> np.ndarray:
        """(r, r, r, 4) f16 bits as clay_volume_download returns them."""
        r = self.res >> m
        out = np.zeros((r, r, r, 4), dtype=np.uint16)
        out[..., 0] = self.level(m, "dist")
        out[..., 1] = self.level(m, "mat")
        out[..., 3] = f32_to_f16(1.0)
        return out


def total_cells(res: int, mips: int) -> int:
    return volume_derive(res, mips)[3]


def edit_bin(res: int, world_size: float, origin, band: float, capacity: int, edits, precision: str = "f32"):
    tr = res // 8
    counts = np.zeros(tr ** 3, dtype=np.uint32)
    ids = np.zeros(tr ** 3 * capacity, dtype=np.uint32)
    getattr(lib(), f"orc_edit_bin_{precision}")(res, C.c_double(world_size), (C.c_double * 3)(*origin),
                                                C.c_double(band), capacity, _brush_array(edits), len(edits),
                                                counts.ctypes.data, ids.ctypes.data)
    return counts, ids.reshape(tr ** 3, capacity)


def tile_dilate(res: int, mips: int, m: int, counts: np.ndarray) -> None:
    assert counts.dtype == np.uint32 and counts.flags.c_contiguous
    lib().orc_tile_dilate(res, mips, m, counts.ctypes.data)


def mip_refresh(planes: Planes, m: int, band: float, safety: float, counts: np.ndarray,
                precision: str = "f32") -> None:
    getattr(lib(), f"orc_mip_refresh_{precision}")(planes.res, planes.mips, m, C.c_double(band),
                                                   C.c_double(safety), counts.ctypes.data,
                                                   planes.dist.ctypes.data, planes.mat.ctypes.data)


def edit_flush(planes: Planes, world_size: float, origin, band: float, edits, safety: float = 0.98,
               capacity: int = 8, precision: str = "f32") -> np.ndarray:
    counts = np.zeros(total_cells(planes.res, planes.mips), dtype=np.uint32)
    getattr(lib(), f"orc_edit_flush_{precision}")(planes.res, planes.mips, C.c_double(world_size),
                                                  (C.c_double * 3)(*origin), C.c_double(band), C.c_double(safety),
                                                  capacity, _brush_array(edits), len(edits),
                                                  planes.dist.ctypes.data, planes.mat.ctypes.data,
                                                  counts.ctypes.data)
    return counts


def contacts_to_edits(raw_count: int, points: np.ndarray, min_speed: float, radius: float, blend: float,
                      max_edits: int = 64) -> list:
    pts = np.zeros((CONTACTS, 4), dtype=np.float32)
    pts[:min(len(points), CONTACTS)] = np.asarray(points, dtype=np.float32)[:CONTACTS]
    arr = (OrcBrush * max_edits)()
    k = lib().orc_contacts_to_edits(raw_count, pts.ctypes.data, C.c_double(min_speed), C.c_double(radius),
                                    C.c_double(blend), max_edits, arr)
    return [arr[i] for i in range(k)]


class OrcExtractParams(C.Structure):
    _fields_ = [
        ("origin", C.c_double * 3), ("cell", C.c_double), ("vel", C.c_double * 3), ("preserve", C.c_int32),
        ("res", C.c_int32), ("body", C.c_uint32), ("base", C.c_uint32), ("capacity", C.c_uint32),
        ("quad_capacity", C.c_uint32),
    ]


class OrcPbdParams(C.Structure):
    _fields_ = [("gravity", C.c_double * 3), ("h", C.c_double), ("radius", C.c_double), ("friction", C.c_double),
                ("damping", C.c_double), ("max_speed", C.c_double), ("push", C.c_double * 3),
                ("kick", C.c_double * 3), ("push_body", C.c_uint32), ("rot_iters", C.c_int32)]


def sin_cos(x: float, precision: str = "f32") -> tuple[float, float]:
    """The oracle's sin and cos (the ClaySolver's quatFromAxisAngle)."""
    dt = C.c_float if precision == "f32" else C.c_double
    s, c = dt(), dt()
    getattr(lib(), f"orc_sin_cos_{precision}")(x, C.byref(s), C.byref(c))
    return s.value, c.value


class Solver:
    """ClaySolver (lib/src/sim/pbd.ts:78-424) over a ParticleSet, in the oracle.

    State is held as numpy arrays: `planes` (15, capacity) = pos, prev, vel,
    velPrev, rest (x, y, z each) and `state` (max_bodies, 10) = com, restCom, rot."""

    def __init__(self, capacity: int, max_bodies: int, gravity=(0.0, -9.81, 0.0), dt: float = 1 / 60,
                 substeps: int = 2, radius: float = 0.1, friction: float = 0.35, damping: float = 0.995,
                 max_speed: float = 40.0, rotation_iterations: int = 2, collider: "Collider | None" = None,
                 precision: str = "f32"):
        self.capacity, self.max_bodies, self.substeps = capacity, max_bodies, substeps
        self.precision = precision
        self.dtype = np.float32 if precision == "f32" else np.float64
        self.k = OrcPbdParams()
        for i in range(3):
            self.k.gravity[i] = gravity[i]
        self.k.h = dt / substeps
        self.k.radius, self.k.friction, self.k.damping, self.k.max_speed = radius, friction, damping, max_speed
        self.k.push_body = 0xFFFFFFFF
        self.k.rot_iters = rotation_iterations
        self.collider = collider
        self.planes = np.zeros((15, capacity), dtype=self.dtype)
        self.aux = np.zeros((4, capacity), dtype=self.dtype)  # restNormal xyz, material
        self.owner = np.full(capacity, 0xFFFFFFFF, dtype=np.uint32)
        self.body3 = np.zeros((max_bodies, 3), dtype=np.uint32)
        self.body2 = np.zeros((max_bodies, 2), dtype=np.float32)
        self.state = np.zeros((max_bodies, 10), dtype=self.dtype)
        self.nbodies = 0
        self.kicking = False

    def set_bodies(self, bodies) -> None:
        """bodies: dicts with first, count, flags, stiffness, plasticity."""
        self.nbodies = len(bodies)
        self.state[:] = 0
        for b, d in enumerate(bodies):
            self.body3[b] = (d["first"], d["count"], d.get("flags", 0))
            self.body2[b] = (d.get("stiffness", 0.5), d.get("plasticity", 0.0))
            self.state[b, 9] = 1.0

    def add_body(self, first: int = 0, count: int = 0, stiffness: float = 0.5, plasticity: float = 0.0,
                 flags: int = 0) -> int:
        """ParticleSet.addBody: one more record, the others untouched."""
        b = self.nbodies
        assert b < self.max_bodies, f"ParticleSet: body limit {self.max_bodies} reached"
        self.body3[b] = (first, count, flags)
        self.body2[b] = (stiffness, plasticity)
        self.state[b] = 0
        self.state[b, 9] = 1.0
        self.nbodies += 1
        return b

    def set_particles(self, pos: np.ndarray, rest: np.ndarray, body: np.ndarray) -> None:
        n = pos.shape[0]
        p = np.asarray(pos, dtype=self.dtype).T
        self.planes[:, :] = 0
        self.aux[:, :] = 0
        self.planes[0:3, :n] = p
        self.planes[3:6, :n] = p
        self.planes[12:15, :n] = np.asarray(rest, dtype=self.dtype).T
        self.owner[:] = 0xFFFFFFFF
        self.owner[:n] = body

    def set_force(self, body: int, accel, impulse=(0.0, 0.0, 0.0)) -> None:
        for i in range(3):
            self.k.push[i] = accel[i]
            self.k.kick[i] = impulse[i]
        self.k.push_body = body
        self.kicking = any(v != 0 for v in impulse)

    def step(self) -> None:
        if self.nbodies == 0:
            return
        col = C.byref(self.collider.c) if self.collider is not None else None
        getattr(lib(), f"orc_pbd_step_{self.precision}")(
            C.byref(self.k), col, self.capacity, self.max_bodies, self.nbodies, self.body3.ctypes.data,
            self.body2.ctypes.data, self.owner.ctypes.data, self.planes.ctypes.data, self.state.ctypes.data,
            1 if self.kicking else 0, self.substeps)
        self.kicking = False

    def get(self, what: int) -> np.ndarray:
        """(capacity, 3) of pos / prev / vel / velPrev / rest (0..4)."""
        return self.planes[3 * what:3 * what + 3].T.copy()


class Extractor:
    """SurfaceExtractor (lib/src/sim/extract.ts:85-484) writing into an oracle Solver's
    particle set; particle slots in ascending cell index, quads in ascending lattice
    point then edge axis."""

    def __init__(self, solver: Solver, collider: "Collider", resolution: int = 32, body: int = 0, base: int = 0,
                 capacity: int = 3000, motion: bool = False):
        assert base + capacity <= solver.capacity, "SurfaceExtractor: slice exceeds set capacity"
        self.s, self.col = solver, collider
        self.k = OrcExtractParams()
        self.k.res, self.k.body, self.k.base, self.k.capacity = resolution, body, base, capacity
        self.k.quad_capacity = int(np.ceil(capacity * 1.5))
        self.k.cell = 0.1
        self.motion = np.zeros(8, dtype=solver.dtype) if motion else None
        self.ids = np.zeros(resolution ** 3, dtype=np.uint32)
        self.indices = np.zeros(6 * max(self.k.quad_capacity, 1), dtype=np.uint32)
        self.counts = np.zeros(2, dtype=np.uint32)
        self.draw = np.zeros(4, dtype=np.uint32)

    def set_region(self, origin, size: float, vel=(0.0, 0.0, 0.0), preserve_motion: bool = False) -> None:
        for i in range(3):
            self.k.origin[i] = origin[i]
            self.k.vel[i] = vel[i]
        self.k.cell = size / self.k.res
        self.k.preserve = 1 if preserve_motion else 0

    def extract(self) -> None:
        s = self.s
        getattr(lib(), f"orc_extract_{s.precision}")(
            C.byref(self.k), C.byref(self.col.c), s.capacity, s.body3.ctypes.data, s.planes.ctypes.data,
            s.aux.ctypes.data, s.owner.ctypes.data, s.state.ctypes.data,
            None if self.motion is None else self.motion.ctypes.data, self.ids.ctypes.data,
            self.indices.ctypes.data, self.counts.ctypes.data, self.draw.ctypes.data)


def body_centroids(first_count, pos: np.ndarray, rest: np.ndarray | None = None,
                   precision: str = "f32") -> np.ndarray:
    fc = np.ascontiguousarray(np.asarray(first_count, dtype=np.uint32).reshape(-1, 2))
    p = np.ascontiguousarray(np.asarray(pos, dtype=np.float32).T)
    r = None if rest is None else np.ascontiguousarray(np.asarray(rest, dtype=np.float32).T)
    out = np.zeros((fc.shape[0], 6), dtype=np.float64)
    rp = [None, None, None] if r is None else [r[0].ctypes.data, r[1].ctypes.data, r[2].ctypes.data]
    getattr(lib(), f"orc_body_centroids_{precision}")(fc.shape[0], fc.ctypes.data, p[0].ctypes.data,
                                                      p[1].ctypes.data, p[2].ctypes.data, *rp, out.ctypes.data)
    return out


# ---------------------------------------------------------------------------
# The baked volume's stream (wire) format, restated in numpy from its definition:
# 96-byte header, a u32 per tile (constant texel or payload index), a bitmap of stored
# tiles, then 4096-byte rgba16float tiles from a 256-aligned offset.  Tiles run
# mip-major, then tile z, y, x; texels within a tile x fastest.
# ---------------------------------------------------------------------------
WIRE_HEADER = np.dtype([
    ("magic", "<u4"), ("version", "<u4"), ("resolution", "<u4"), ("mip_levels", "<u4"),
    ("band", "<f4"), ("world_size", "<f4"), ("gradient_voxels", "<f4"), ("tiles", "<u4"),
    ("origin", "<f8", 3), ("stored_tiles", "<u4"), ("reserved", "<u4"),
    ("table_offset", "<u8"), ("bitmap_offset", "<u8"), ("payload_offset", "<u8"), ("total_bytes", "<u8"),
])
assert WIRE_HEADER.itemsize == 96


def _tiles(a: np.ndarray) -> np.ndarray:
    """(r, r, r) [z, y, x] -> (tiles, 512) in tile z, y, x order, texels x fastest."""
    t = a.shape[0] // 8
    return a.reshape(t, 8, t, 8, t, 8).transpose(0, 2, 4, 1, 3, 5).reshape(t ** 3, 512)


def _untiles(b: np.ndarray, r: int) -> np.ndarray:
    t = r // 8
    return b.reshape(t, t, t, 8, 8, 8).transpose(0, 3, 1, 4, 2, 5).reshape(r, r, r)


def wire_layout(tiles: int):
    table = 96
    bitmap = table + 4 * tiles
    payload = (bitmap + 4 * ((tiles + 31) // 32) + 255) & ~255
    return table, bitmap, payload


def wire_pack(dist, mat, band: float, world_size: float, origin, gradient_voxels: float = 1.0) -> bytes:
    """dist, mat: per-mip (r, r, r) uint16 planes (f16 bits) indexed [z, y, x]."""
    d = np.concatenate([_tiles(np.asarray(x, np.uint16)) for x in dist])
    m = np.concatenate([_tiles(np.asarray(x, np.uint16)) for x in mat])
    tiles = d.shape[0]
    const = (d == d[:, :1]).all(1) & (m == m[:, :1]).all(1)
    table = np.where(const, (m[:, 0].astype(np.uint32) << 16) | d[:, 0], 0).astype(np.uint32)
    stored = np.flatnonzero(~const)
    table[stored] = np.arange(stored.size, dtype=np.uint32)
    bits = np.zeros(((tiles + 31) // 32) * 32, dtype=np.uint8)
    bits[stored] = 1
    bitmap = np.packbits(bits.reshape(-1, 32)[:, ::-1], axis=1).view(">u4").astype("<u4").ravel()
    t_off, b_off, p_off = wire_layout(tiles)
    texels = np.zeros((stored.size, 512, 4), dtype=np.uint16)
    texels[..., 0] = d[stored]
    texels[..., 1] = m[stored]
    texels[..., 3] = 0x3C00
    h = np.zeros((), dtype=WIRE_HEADER)
    h["magic"], h["version"] = 0x42564C43, 1
    h["resolution"], h["mip_levels"] = dist[0].shape[0], len(dist)
    h["band"], h["world_size"], h["gradient_voxels"], h["tiles"] = band, world_size, gradient_voxels, tiles
    h["origin"] = origin
    h["stored_tiles"] = stored.size
    h["table_offset"], h["bitmap_offset"], h["payload_offset"] = t_off, b_off, p_off
    h["total_bytes"] = p_off + 4096 * stored.size
    out = bytearray(int(h["total_bytes"]))
    out[0:96] = h.tobytes()
    out[t_off:t_off + 4 * tiles] = table.astype("<u4").tobytes()
    out[b_off:b_off + bitmap.nbytes] = bitmap.tobytes()
    out[p_off:] = texels.astype("<u2").tobytes()
    return bytes(out)


def wire_unpack(data: bytes):
    """Returns (header record, per-mip dist planes, per-mip mat planes)."""
    h = np.frombuffer(data[:96], dtype=WIRE_HEADER)[0]
    if h["magic"] != 0x42564C43 or h["version"] != 1:
        raise ValueError("volume stream: bad magic or version")
    if len(data) < h["total_bytes"]:
        raise ValueError("volume stream: truncated")
    tiles = int(h["tiles"])
    table = np.frombuffer(data, dtype="<u4", count=tiles, offset=int(h["table_offset"]))
    words = np.frombuffer(data, dtype="<u4", count=(tiles + 31) // 32, offset=int(h["bitmap_offset"]))
    stored = ((words[:, None] >> np.arange(32, dtype=np.uint32)) & 1).ravel()[:tiles].astype(bool)
    n = int(h["stored_tiles"])
    pay = np.frombuffer(data, dtype="<u2", count=n * 2048, offset=int(h["payload_offset"])).reshape(n, 512, 4)
    if (table[stored] >= n).any():
        raise ValueError("volume stream: a stored tile's index lies outside the payload")
    d = np.where(stored[:, None], 0, (table & 0xFFFF)[:, None]).astype(np.uint16) * np.ones((1, 512), np.uint16)
    m = np.where(stored[:, None], 0, (table >> 16)[:, None]).astype(np.uint16) * np.ones((1, 512), np.uint16)
    d[stored] = pay[table[stored], :, 0]
    m[stored] = pay[table[stored], :, 1]
    dist, mat, at = [], [], 0
    for lvl in range(int(h["mip_levels"])):
        r = int(h["resolution"]) >> lvl
        k = (r // 8) ** 3
        dist.append(_untiles(d[at:at + k], r))
        mat.append(_untiles(m[at:at + k], r))
        at += k
    return h, dist, mat
