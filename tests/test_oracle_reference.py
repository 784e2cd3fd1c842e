"""The oracle against the reference's own known answers and against mathematics.

Ports of the reference's JS mirror tests (lib/test/fluid.test.mjs:30-145,
lib/test/march.test.mjs:176-182), its GPU tests' closed-form answers
(lib/test/gpu/volume.gpu.ts:28-67, lib/test/gpu/brush.gpu.ts:34-91), plus
brute-force and invariant checks chosen so that a dropped term, a wrong sign or
index, or a transposed operand fails one of them.
"""
import math

import numpy as np
import pytest

import oracle as O
from clayhip import make_brush, scenes


# --- the reference's seeded LCG, with JS number semantics (fluid.test.mjs:15-19) ---
class Lcg:
    def __init__(self, seed: int = 12345):
        self.seed = seed

    def __call__(self) -> float:
        # JS: (seed * 1103515245 + 12345) & 0x7fffffff, products rounded in f64
        v = float(self.seed) * 1103515245.0 + 12345.0
        self.seed = int(v) & 0x7FFFFFFF
        return self.seed / 0x7FFFFFFF

    def cloud(self, n, lo, hi):
        return [[lo + self() * (hi - lo) for _ in range(3)] for _ in range(n)]


def test_lcg_port_matches_js_int32_semantics():
    # First values of the reference LCG computed by hand with exact integer arithmetic
    # where the f64 product is still exact (seed < 2^22 keeps it below 2^53).
    r = Lcg(1)
    assert r() == ((1103515245 + 12345) & 0x7FFFFFFF) / 0x7FFFFFFF


def test_three_by_three_scan_is_a_complete_neighbour_search(oracle):
    """fluid.test.mjs:30-65 on the oracle's grid and bucket scan."""
    rnd = Lcg()
    pts = np.array(rnd.cloud(400, -3, 3), dtype=np.float64)
    opts = {"spacing": 0.45, "smoothing": 1.0, "origin": (-12, -4, -12), "worldSize": 24.0}
    k = O.fluid_consts(opts)
    assert k.grid_res == 26 and tuple(k.grid_origin) == (-13.0, -5.0, -13.0)
    # Dense enough that most cells are occupied (the reference's own guard).
    keys = O.keys(opts, pts.astype(np.float32))
    assert (keys != O.INVALID_KEY).all(), "grid must cover the volume"
    d = np.sqrt(((pts[:, None, :] - pts[None, :, :]) ** 2).sum(-1))
    for i in range(len(pts)):
        brute = sorted(np.nonzero(d[i] < 1.0)[0].tolist())
        assert O.neighbors(opts, pts, i) == brute, f"particle {i}"


def test_splat_bounds_cover_every_voxel_the_particle_can_lower(oracle):
    """fluid.test.mjs:67-91: with one particle at a time, every voxel whose exact
    distance is inside the band got a code, and no voxel outside the band did."""
    res, ws = 32, 24.0
    voxel = ws / res
    surf_r = 0.5
    rnd = Lcg()
    centres = (np.indices((res, res, res)).transpose(1, 2, 3, 0)[..., ::-1] + 0.5) * voxel  # [z,y,x] -> xyz
    for band in (2, 4):
        bw = band * voxel
        o = O.splat_opts(res, 1, ws, (0, 0, 0), band, surf_r)
        for p in rnd.cloud(60, 1, 22):
            g = O.splat_bake(o, np.array([p], dtype=np.float32))[0]
            sd = np.sqrt(((centres - np.array(p)) ** 2).sum(-1)) - surf_r
            inside = sd < bw - 1e-4
            outside = sd >= bw + 1e-4
            assert (g[inside] < 65535).all(), f"voxel inside the band missed for band {band}"
            assert (g[outside] == 65535).all()


def test_splat_equals_brute_force_gather(oracle):
    """The scatter (box walk + atomicMin) equals a per-voxel gather over all particles."""
    rng = np.random.default_rng(4)
    pos = rng.uniform(0.5, 5.5, size=(40, 3)).astype(np.float32)
    res, ws, band, rad = 16, 6.0, 3, 0.4
    o = O.splat_opts(res, 2, ws, (0, 0, 0), band, rad)
    grids = O.splat_bake(o, pos)
    for m in range(2):
        r = res >> m
        v = np.float32(ws / res * 2 ** m)
        bw = np.float32(band * ws / res * 2 ** m)
        idx = np.indices((r, r, r)).transpose(1, 2, 3, 0)[..., ::-1].astype(np.float32)
        c = (idx + np.float32(0.5)) * v
        best = np.full((r, r, r), 65535, dtype=np.uint32)
        for p in pos:
            dd = c - p
            sd = np.sqrt(dd[..., 0] * dd[..., 0] + dd[..., 1] * dd[..., 1] + dd[..., 2] * dd[..., 2]) - np.float32(rad)
            q = (np.clip(sd / bw, np.float32(-1), np.float32(1)) + np.float32(1)) * np.float32(32767.5)
            e = np.where(sd < bw, q.astype(np.uint32), np.uint32(65535))
            best = np.minimum(best, e)
        assert np.array_equal(grids[m], best), f"mip {m}"


def test_encoding_turns_atomic_min_into_a_distance_union(oracle):
    """fluid.test.mjs:93-113."""
    assert O.enc(-1.0) == 0
    assert O.enc(1.0) == 65535
    prev = -1
    for i in range(401):
        e = O.enc(-1.5 + i * 0.0075)
        assert e >= prev, "encoding must be non-decreasing"
        prev = e
    rnd = Lcg()
    for _ in range(500):
        a, b = rnd() * 4 - 2, rnd() * 4 - 2
        assert min(O.enc(a), O.enc(b)) == O.enc(min(a, b))
    for _ in range(500):
        v = rnd() * 2 - 1
        assert abs(O.dec(O.enc(v)) - v) <= 1 / 32767.5


def test_rest_density_is_the_density_of_its_lattice(oracle):
    """fluid.test.mjs:115-145."""
    spacing = 0.45
    k = O.fluid_consts({"spacing": spacing})
    h = spacing * 2.2
    assert k.h == h
    poly6 = 315 / (64 * math.pi * h ** 9)
    mass = spacing ** 3
    measured = 0.0
    for z in range(-6, 7):
        for y in range(-6, 7):
            for x in range(-6, 7):
                r2 = (x * spacing) ** 2 + (y * spacing) ** 2 + (z * spacing) ** 2
                if r2 < h * h:
                    measured += mass * poly6 * (h * h - r2) ** 3
    assert abs(measured - k.rest_density) < 1e-9
    assert 0.5 < k.rest_density < 2


def test_kernel_constants_are_muller_2003(oracle):
    """poly6 = 315/(64 pi h^9), spiky gradient = -45/(pi h^6) (Mueller et al. 2003);
    poly6 integrates to one over the ball (checked numerically)."""
    k = O.fluid_consts({"spacing": 0.05, "smoothing": 0.1})
    h = 0.1
    assert math.isclose(k.poly6, 315 / (64 * math.pi * h ** 9), rel_tol=1e-15)
    assert math.isclose(k.spiky_grad, -45 / (math.pi * h ** 6), rel_tol=1e-15)
    r = np.linspace(0, h, 200001)
    integral = np.trapezoid(k.poly6 * (h * h - r * r) ** 3 * 4 * math.pi * r * r, r)
    assert abs(integral - 1) < 1e-6
    assert math.isclose(k.dt, 1 / 60 / 3) and k.grid_res == math.ceil(24 / h) + 2


def test_pack_tile_round_trips_up_to_1023(oracle):
    """march.test.mjs:176-182."""
    for c in [(0, 0, 0), (1, 2, 3), (1023, 1023, 1023), (7, 0, 511)]:
        assert O.unpack_tile(O.pack_tile(*c)) == c


def test_volume_mip_derivation_and_refusals(oracle):
    """volume.gpu.ts:28-67."""
    m, tr, co, tot = O.volume_derive(160)
    assert (m, tr, tot) == (3, [20, 10, 5], 20 ** 3 + 10 ** 3 + 5 ** 3)
    assert co == [0, 8000, 9000]
    assert O.volume_derive(40)[:2] == (1, [5])
    for res, levels in [(64, 3), (128, 4), (256, 5)]:
        assert O.volume_derive(res)[0] == levels
    with pytest.raises(ValueError, match="must be a multiple of 64"):
        O.volume_derive(160, 4)
    with pytest.raises(ValueError, match="multiple of 8"):
        O.volume_derive(100)
    with pytest.raises(ValueError, match="resolution 288 must be a multiple of 128 for 5 mip levels"):
        O.volume_derive(288, 5)
    assert O.volume_derive(288, 3)[0] == 3


def _eval(brush, pts):
    col = O.Collider().add_analytic([make_brush(brush)])
    d, _ = col.eval(np.array(pts, dtype=np.float32), precision="f64")
    return d


def test_primitive_distances_match_closed_forms(oracle):
    """brush.gpu.ts:34-91 (tolerance 2e-3 as the reference's `near`)."""
    near = lambda a, b: abs(a - b) < 2e-3  # noqa: E731
    d = _eval({"kind": "sphere", "size": 2, "pos": (1, 0, 0)}, [[1, 0, 0], [4, 0, 0], [3, 0, 0], [1, 0, 5]])
    assert all(near(a, b) for a, b in zip(d, [-2, 1, 0, 3]))
    d = _eval({"kind": "sphere", "size": 1, "scale": 3}, [[6, 0, 0], [0, 0, 0]])
    assert near(d[0], 3) and near(d[1], -3)
    d = _eval({"kind": "capsule", "size": (0, 2, 0), "radius": 0.5, "euler": (0, 0, math.pi / 2)},
              [[1, 0, 0], [0, 1, 0], [3, 0, 0]])
    assert near(d[0], -0.5) and near(d[1], 0.5) and near(d[2], 0.5)
    d = _eval({"kind": "plane", "radius": 1.5}, [[0, 3, 0], [9, 0, -9]])
    assert near(d[0], 1.5) and near(d[1], -1.5)
    d = _eval({"kind": "box", "size": (1, 2, 3)}, [[2, 0, 0], [0, 0, 0], [1, 2, 3]])
    assert near(d[0], 1) and near(d[1], -1) and near(d[2], 0)


def test_round_box_is_the_quilez_form(oracle):
    """sdRoundedBox3d: parity unpinned against @typegpu/sdf (no reference test);
    pinned here to Quilez's closed form: size is the outer half-extent, corners
    rounded by r, so faces sit at +-size and an edge is r*(sqrt(2)-1) inside."""
    b = {"kind": "roundBox", "size": (11, 1, 11), "radius": 0.25, "pos": (0, -1.4, 0)}
    d = _eval(b, [[0, 0.6, 0], [0, -0.4, 0], [0, -1.4, 0], [11.5, -1.4, 0], [11 + 1, 0, 0]])
    assert abs(d[0] - 1.0) < 1e-6 and abs(d[1]) < 1e-6 and abs(d[2] + 1.0) < 1e-6
    assert abs(d[3] - 0.5) < 1e-6
    # Diagonally off the (x, y) edge: distance to the rounded edge's arc, whose
    # centre line runs at (11 - r, -0.4 - r).
    expect = math.hypot(12.0 - (11 - 0.25), 0.0 - (-0.4 - 0.25)) - 0.25
    assert abs(d[4] - expect) < 1e-6


def test_hash1_is_lowbias32(oracle):
    """math/gpu.ts:103-112 is Wellons' lowbias32 (constants 0x7feb352d, 0x846ca68b);
    every step is invertible, so the hash is a bijection: check by inverting it."""
    def inv(x):
        x ^= x >> 16
        x = (x * 0x43021123) & 0xFFFFFFFF  # inverse of 0x846ca68b mod 2^32
        x ^= (x >> 15) ^ (x >> 30)
        x = (x * 0x1D69E2A5) & 0xFFFFFFFF  # inverse of 0x7feb352d mod 2^32
        x ^= x >> 16
        return x
    assert (0x846CA68B * 0x43021123) & 0xFFFFFFFF == 1
    assert (0x7FEB352D * 0x1D69E2A5) & 0xFFFFFFFF == 1
    for s in [0, 1, 2, 12345, 0xDEADBEEF, 0xFFFFFFFF] + list(range(100, 5000, 97)):
        assert inv(O.hash_u32(s)) == s
        assert 0.0 <= O.hash1(s) < 1.0
    assert np.array_equal(scenes.hash_u32(np.arange(1000, dtype=np.uint32)),
                          np.array([O.hash_u32(i) for i in range(1000)], dtype=np.uint32))


def test_f16_conversion_matches_numpy(oracle):
    rng = np.random.default_rng(0)
    vals = np.concatenate([
        rng.normal(size=4000).astype(np.float32),
        rng.uniform(-70000, 70000, 2000).astype(np.float32),
        (rng.uniform(-1, 1, 2000) * 2.0 ** -20).astype(np.float32),
        np.array([0.0, -0.0, 1.0, -1.0, 65504.0, 65520.0, 1e-8, 2.0 ** -24, 2.0 ** -25, 3 * 2.0 ** -26], np.float32),
    ])
    # exact halfway cases between adjacent f16 values
    h = np.arange(0, 0x7BFF, 7, dtype=np.uint16).view(np.float16).astype(np.float32)
    h2 = np.arange(1, 0x7C00, 7, dtype=np.uint16).view(np.float16).astype(np.float32)
    n = min(len(h), len(h2))
    vals = np.concatenate([vals, ((h[:n].astype(np.float64) + h2[:n]) / 2).astype(np.float32)])
    got = np.array([O.f32_to_f16(float(v)) for v in vals], dtype=np.uint16)
    want = vals.astype(np.float16).view(np.uint16)
    assert np.array_equal(got, want)
    back = np.array([O.f16_to_f32(int(u)) for u in want], dtype=np.float32)
    assert np.array_equal(back, want.view(np.float16).astype(np.float32))


def test_stable_sort_matches_numpy(oracle):
    rng = np.random.default_rng(1)
    for n in (0, 1, 7, 1000, 70000):
        k = rng.integers(0, 300, size=n).astype(np.uint32)
        k[::13] = O.INVALID_KEY
        assert np.array_equal(O.stable_sort(k), np.argsort(k, kind="stable").astype(np.uint32))


def test_lattice_at_rest_has_rest_density_and_no_pressure_force(oracle):
    """An unjittered lattice far from walls and colliders: interior particles read
    rho = rho0 and feel only gravity (nvel = g dt damping)."""
    spacing, h = 0.05, 0.1
    pos = scenes.lattice(14, 14, 14, spacing, (0.0, 2.0, 0.0), jitter=0.0)
    opts = {"spacing": spacing, "smoothing": h, "density": 1000.0, "substeps": 2}
    fl = O.Fluid(opts, pos, precision="f64")
    fl.substep()
    rho = fl.get("density")
    ijk = np.stack(np.meshgrid(np.arange(14), np.arange(14), np.arange(14), indexing="ij"), -1).reshape(-1, 3)[:, ::-1]
    interior = np.all((ijk >= 3) & (ijk <= 10), axis=1)
    # Positions are f32-rounded lattice points, so rho matches rho0 to ~1e-6; the
    # self term alone is 19% of rho0 and one missing neighbour ring is > 1%.
    assert np.allclose(rho[interior], fl.k.rest_density, rtol=1e-5)
    assert (rho[~interior] < fl.k.rest_density).any()
    nv = fl.get("nvel")[interior]
    want_vy = -9.81 * fl.k.dt * fl.k.damping
    # The f32-rounded lattice leaves pressures of order 0.1 whose gradients almost
    # cancel: gravity dominates to 1e-3, sideways forces stay below 1e-4 m/s.
    assert np.allclose(nv[:, 1], want_vy, rtol=1e-3, atol=1e-9)
    assert np.abs(nv[:, [0, 2]]).max() < 1e-4


def test_f32_and_f64_oracles_agree_per_substep(oracle):
    sc = scenes.dam_break_small()
    col = O.Collider().add_analytic([make_brush(b) for b in sc["collider"]["brushes"]])
    a = O.Fluid(sc["options"], sc["pos"], collider=col, precision="f32")
    b = O.Fluid(sc["options"], sc["pos"], collider=col, precision="f64")
    a.substep()
    b.substep()
    rho_a, rho_b = a.get("density"), b.get("density")
    assert np.max(np.abs(rho_a - rho_b) / rho_b) < 1e-5
    assert np.max(np.abs(a.get("nvel") - b.get("nvel"))) < 1e-3
    assert np.max(np.abs(a.get("pos") - b.get("pos"))) < 1e-5


def test_set_origin_by_whole_voxel_shifts_the_grid(oracle):
    rng = np.random.default_rng(3)
    pos = rng.uniform(2, 4, size=(50, 3)).astype(np.float32)
    res, ws = 32, 8.0
    v = ws / res
    g0 = O.splat_bake(O.splat_opts(res, 1, ws, (0, 0, 0), 3, 0.2), pos)[0]
    g1 = O.splat_bake(O.splat_opts(res, 1, ws, (v, 0, 0), 3, 0.2), pos)[0]
    # Voxel x+1 at origin 0 is voxel x at origin v (same centre in world space);
    # the two evaluations round differently only in the last bit of the distance.
    a, b = g1[:, :, :-1].astype(np.int64), g0[:, :, 1:].astype(np.int64)
    both = (a < 65535) & (b < 65535)
    assert both.sum() > 1000
    assert np.abs(a[both] - b[both]).max() <= 1
    assert (a == b).mean() > 0.99


def test_brush_bake_saturates_away_from_brushes(oracle):
    recs = [make_brush(b) for b in scenes.ridge_world_brushes()]
    tex = O.bake_brushes(64, 3, 24.0, (-12, -4, -12), 4.0, recs)
    d0 = tex[0][..., 0].view(np.float16).astype(np.float32)
    # Band 4 voxels of 0.375 = 1.5 world units: the 1-unit-thick floor half and the
    # 0.9 ridge half never saturate inside; far from every brush the field is +1.
    assert d0.max() == 1.0 and -0.7 < d0.min() < -0.5
    v = 24 / 64
    ix = iz = int(12 / v)  # x = z = 0 (+ half a voxel)
    floor_mid = int((-1.4 + 4) / v)
    above = int((5.0 + 4) / v)
    assert d0[iz, floor_mid, ix] < -0.5 and d0[iz, above, ix] == 1.0
    mat = tex[0][..., 1].view(np.float16)
    assert (tex[0][..., 3].view(np.float16) == 1).all() and (mat == 0).all()
