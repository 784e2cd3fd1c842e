"""The oracle's density, force and integrate passes and its collider normals,
against an independent numpy restatement and against closed forms.

The HIP kernels equal the oracle bit for bit, so these checks are what ties both
to the reference (lib/src/sim/fluid.ts:333-491, lib/src/trace/field.ts:225-234,
lib/src/field/volume.ts:299-313).  Each is chosen so that a plausible transcription
mistake fails it: the sign of the pressure term, a dropped 0.5 or (p_i + p_j),
(h - r) in place of (h - r)^2, the XSPH seed or blend, the speed cap, the collision
push-out and velocity kill, the contact threshold, the wall clamp, and the normals.
"""
import math

import numpy as np
import pytest

import oracle as O
import sph_numpy as R
from clayhip import make_brush

FLOOR = {"kind": "roundBox", "size": (11, 1, 11), "radius": 0.25, "pos": (0, -1.4, 0)}
FLOOR_TOP = float(np.float32(-1.4)) + 1.0  # the brush record holds f32 values


def floor_sdf(p):
    return R.round_box(p, (0.0, float(np.float32(-1.4)), 0.0), (11.0, 1.0, 11.0), 0.25)


def floor_normal(p):
    return R.central_normal(floor_sdf, p, 1e-3)


def floor_collider():
    return O.Collider().add_analytic([make_brush(FLOOR)])


def opts(**kw):
    o = {"spacing": 0.05, "smoothing": 0.1, "density": 1000.0, "substeps": 2, "origin": (-1.0, -1.0, -1.0),
         "worldSize": 2.0, "gravity": (0.0, -9.81, 0.0)}
    o.update(kw)
    return o


def params(fl):
    k = fl.k
    return dict(gravity=tuple(k.gravity), viscosity=k.viscosity, damping=k.damping, max_speed=k.max_speed)


def consts(fl):
    k = fl.k
    return {"h": k.h, "poly6": k.poly6, "spiky": k.spiky_grad, "mass": k.mass, "rest": k.rest_density,
            "dt": k.dt}


def moving_block(seed=1):
    """A compressed, jittered block with random velocities, resting on the floor top
    (y = -0.4) and pressed against the +x wall (x = 1 - radius): every branch of the
    passes is taken (positive pressures, neighbours beyond h, collisions, contacts,
    clamps, speeds above the cap)."""
    rng = np.random.default_rng(seed)
    g = np.stack(np.meshgrid(np.arange(12), np.arange(10), np.arange(10), indexing="ij"), -1).reshape(-1, 3)
    pos = np.array([0.55, FLOOR_TOP + 0.012, -0.2]) + g * 0.041 + rng.uniform(-0.008, 0.008, size=g.shape)
    vel = rng.normal(0.0, 1.5, size=pos.shape)
    vel[:, 1] -= 2.0
    vel[:, 0] += 1.0
    vel[::37] *= 12.0  # a few above the 15 m/s cap
    return pos, vel


# ---- constants -------------------------------------------------------------
def test_constants_match_the_reference_formulas(oracle):
    fl = O.Fluid(opts(), np.zeros((1, 3)), precision="f64")
    want = R.constants(0.05, 0.1, density=1000.0, substeps=2)
    got = consts(fl)
    for key in ("poly6", "spiky", "mass", "rest", "dt"):
        assert got[key] == pytest.approx(want[key], rel=1e-14), key
    k = fl.k
    assert k.contact_depth == pytest.approx(1.5 * 0.025, rel=1e-15)
    assert tuple(k.wall_lo) == pytest.approx((-1 + 0.025,) * 3) and tuple(k.wall_hi) == pytest.approx((1 - 0.025,) * 3)
    # poly6 integrates to 1 over the ball of radius h (Mueller et al. 2003)
    h = 0.1
    r = np.linspace(0, h, 200001)
    f = want["poly6"] * (h * h - r * r) ** 3 * 4 * math.pi * r * r
    assert np.sum((f[1:] + f[:-1]) * 0.5 * np.diff(r)) == pytest.approx(1.0, rel=1e-9)


# ---- independent restatement, per pass and per substep -----------------------
def test_passes_equal_the_numpy_restatement_per_substep(oracle):
    pos, vel = moving_block()
    fl = O.Fluid(opts(), pos, vel, collider=floor_collider(), precision="f64")
    k = consts(fl)
    prm = params(fl)
    for sub in range(4):
        p0, v0 = fl.get("pos").copy(), fl.get("vel").copy()
        fl.density()
        rho, pres = R.density(p0, k, fl.k.stiffness)
        assert np.allclose(fl.get("density"), rho, rtol=1e-12, atol=0), f"substep {sub}: density"
        assert np.allclose(fl.get("pressure"), pres, rtol=1e-10, atol=1e-9), f"substep {sub}: pressure"
        assert (pres > 0).sum() > 100, "the block must be compressed"
        fl.forces()
        nv = R.forces(p0, v0, fl.get("density"), fl.get("pressure"), k, **prm)
        assert np.allclose(fl.get("nvel"), nv, rtol=1e-10, atol=1e-12), f"substep {sub}: nvel"
        assert (np.linalg.norm(v0, axis=1) > 15).any()
        nvel = fl.get("nvel").copy()
        fl.integrate()
        want_p, want_v, ids, pts = R.integrate(p0, nvel, k["dt"], fl.k.radius, fl.k.contact_depth, floor_sdf,
                                               floor_normal, fl.k.wall_lo, fl.k.wall_hi)
        assert np.allclose(fl.get("pos"), want_p, rtol=0, atol=1e-13), f"substep {sub}: pos"
        assert np.allclose(fl.get("vel"), want_v, rtol=0, atol=1e-12), f"substep {sub}: vel"
        raw, got_pts = fl.read_contacts()
        assert raw == len(ids) and raw > 0, f"substep {sub}: contact count"
        assert np.allclose(np.array(got_pts), pts[:O.CONTACTS], rtol=0, atol=1e-12), f"substep {sub}: contacts"
        assert (fl.get("pos")[:, 0] == fl.k.wall_hi[0]).any(), "the wall must clamp someone"


def test_numpy_restatement_catches_planted_mistakes(oracle):
    """The comparison above is not vacuous: each plausible mistake, planted in the
    numpy side, moves nvel far beyond its tolerance."""
    pos, vel = moving_block(seed=2)
    fl = O.Fluid(opts(), pos, vel, collider=floor_collider(), precision="f64")
    fl.density()
    fl.forces()
    got = fl.get("nvel")
    k = consts(fl)
    prm = params(fl)
    rho, pres = fl.get("density"), fl.get("pressure")
    mistakes = {
        "pressure sign": dict(k, spiky=-k["spiky"]),
        "dropped 0.5": dict(k, spiky=2 * k["spiky"]),
    }
    for name, kk in mistakes.items():
        bad = R.forces(pos, vel, rho, pres, kk, **prm)
        assert not np.allclose(got, bad, rtol=1e-10, atol=1e-12), name
    bad = R.forces(pos, vel, rho, pres * 0 + pres.mean(), k, **prm)  # p_i + p_j replaced
    assert not np.allclose(got, bad, rtol=1e-10, atol=1e-12), "pressure pairing"
    bad = R.forces(pos, vel, rho, pres, k, **dict(prm, viscosity=0.0))
    assert not np.allclose(got, bad, rtol=1e-10, atol=1e-12), "XSPH blend"


# ---- closed forms: pressure ------------------------------------------------------
def test_pressure_forces_cancel_in_total(oracle):
    """With gravity, viscosity and damping off and zero velocities, nvel_i is
    dt fPress_i / rho_i, and the pair term m^2 (p_i + p_j) / (2 rho_i rho_j) grad W
    is antisymmetric: the total m sum(nvel) vanishes to rounding."""
    pos, _ = moving_block(seed=3)
    fl = O.Fluid(opts(gravity=(0.0, 0.0, 0.0), viscosity=0.0, damping=1.0, maxSpeed=1e9), pos,
                 precision="f64")
    fl.density()
    fl.forces()
    nv = fl.get("nvel")
    scale = np.abs(nv).sum(0)
    assert (scale > 1.0).all(), "the block must push outwards"
    assert np.all(np.abs(nv.sum(0)) < 1e-12 * scale)


@pytest.mark.parametrize("r", [0.02, 0.05, 0.09])
def test_two_compressed_particles_fly_apart_with_the_spiky_gradient(oracle, r):
    """Spacing 0.15 > h makes rho0 the self term alone, so two particles r < h apart
    have pressure k m poly6 (h^2 - r^2)^3 each and accelerate apart along the line
    between them with |a| = m p / rho^2 * 45 / (pi h^6) (h - r)^2."""
    h = 0.1
    pos = np.array([[0.0, 0.3, 0.0], [r, 0.3, 0.0]])
    fl = O.Fluid(opts(spacing=0.15, density=1.0, gravity=(0.0, 0.0, 0.0), viscosity=0.0, damping=1.0,
                      maxSpeed=1e9), pos, precision="f64")
    fl.density()
    fl.forces()
    k = R.constants(0.15, h, density=1.0, substeps=2)
    assert fl.k.rest_density == pytest.approx(k["mass"] * k["poly6"] * h ** 6, rel=1e-14)
    rho = k["mass"] * k["poly6"] * (h ** 6 + (h * h - r * r) ** 3)
    p = 220.0 * (rho - fl.k.rest_density)
    assert fl.get("density") == pytest.approx([rho, rho], rel=1e-13)
    assert fl.get("pressure") == pytest.approx([p, p], rel=1e-12)
    a = k["mass"] * p / (rho * rho) * 45.0 / (math.pi * h ** 6) * (h - r) ** 2
    nv = fl.get("nvel")
    assert nv[0, 0] == pytest.approx(-a * k["dt"], rel=1e-12)
    assert nv[1, 0] == pytest.approx(a * k["dt"], rel=1e-12)
    assert np.all(nv[:, 1:] == 0)


# ---- closed forms: XSPH and the cap ------------------------------------------------
def test_xsph_keeps_uniform_and_lone_velocities(oracle):
    """A uniform velocity field is a fixed point of the blend, and so is a lone
    particle (its own weight h^6 seeds the sums)."""
    pos, _ = moving_block(seed=4)
    pos[-1] = (0.0, 0.5, 0.0)  # far from the block: alone
    vel = np.tile([0.7, -0.3, 0.2], (pos.shape[0], 1))
    vel[-1] = (-2.0, 1.0, 3.0)
    fl = O.Fluid(opts(gravity=(0.0, 0.0, 0.0), viscosity=1.0, damping=1.0, maxSpeed=1e9, stiffness=0.0), pos, vel,
                 precision="f64")
    fl.density()
    fl.forces()
    nv = fl.get("nvel")
    assert np.allclose(nv[:-1], vel[:-1], rtol=0, atol=1e-15)
    assert np.array_equal(nv[-1], vel[-1])


def test_xsph_is_a_convex_blend_of_the_neighbourhood(oracle):
    """With viscosity 1 every component of nvel lies between the smallest and the
    largest of that component over the particle and its neighbours within h; with
    viscosity t it is mix(v_i, that blend, t)."""
    pos, vel = moving_block(seed=5)
    base = opts(gravity=(0.0, 0.0, 0.0), damping=1.0, maxSpeed=1e9, stiffness=0.0)
    full = O.Fluid(dict(base, viscosity=1.0), pos, vel, precision="f64")
    full.density()
    full.forces()
    blend = full.get("nvel")
    d = np.sqrt(((pos[:, None, :] - pos[None, :, :]) ** 2).sum(-1))
    near = d < 0.1
    for a in range(3):
        lo = np.where(near, vel[None, :, a], np.inf).min(1)
        hi = np.where(near, vel[None, :, a], -np.inf).max(1)
        assert np.all(blend[:, a] >= lo - 1e-12) and np.all(blend[:, a] <= hi + 1e-12)
    part = O.Fluid(dict(base, viscosity=0.4), pos, vel, precision="f64")
    part.density()
    part.forces()
    assert np.allclose(part.get("nvel"), vel * 0.6 + blend * 0.4, rtol=0, atol=1e-12)


def test_speed_cap(oracle):
    pos, vel = moving_block(seed=6)
    fl = O.Fluid(opts(), pos, vel, precision="f64")
    fl.density()
    fl.forces()
    s = np.linalg.norm(fl.get("nvel"), axis=1)
    assert s.max() <= 15.0 * (1 + 1e-15) and np.isclose(s, 15.0, rtol=1e-14).sum() >= 5
    free = O.Fluid(opts(maxSpeed=1e9), pos, vel, precision="f64")
    free.density()
    free.forces()
    s0 = np.linalg.norm(free.get("nvel"), axis=1)
    slow = s0 <= 15.0
    assert np.array_equal(fl.get("nvel")[slow], free.get("nvel")[slow])
    fast = ~slow
    want = free.get("nvel")[fast] * (15.0 / s0[fast])[:, None]
    assert np.allclose(fl.get("nvel")[fast], want, rtol=1e-14)


# ---- closed forms: integrate -------------------------------------------------------
def _integrate_one(p, nv, **kw):
    fl = O.Fluid(opts(**kw), np.array([p], dtype=np.float64), collider=floor_collider(), precision="f64")
    fl.set("nvel", np.array([nv], dtype=np.float64))
    fl.integrate()
    return fl, fl.get("pos")[0], fl.get("vel")[0], fl.read_contacts()


def test_collision_pushes_out_to_the_radius_and_kills_only_inward_velocity(oracle):
    dt, rad = 1.0 / 120.0, 0.025
    # Over the flat top, the floor's distance is y - top and its normal is +y.
    for y0, vy in ((0.01, -3.0), (0.02, -1.0), (-0.004, -0.5)):
        p = (0.3, FLOOR_TOP + y0, -0.2)
        fl, pos, vel, (raw, pts) = _integrate_one(p, (1.0, vy, 0.5))
        dw = (p[1] + vy * dt) - FLOOR_TOP
        if dw < rad:
            assert pos[1] == pytest.approx(FLOOR_TOP + rad, abs=1e-12)
            assert vel[1] == 0 and vel[0] == 1.0 and vel[2] == 0.5
            assert pos[0] == pytest.approx(0.3 + dt) and pos[2] == pytest.approx(-0.2 + 0.5 * dt)
            assert raw == 1 and pts[0] == pytest.approx((pos[0], pos[1], pos[2], math.hypot(1.0, 0.5)))
        else:
            assert pos[1] == pytest.approx(p[1] + vy * dt) and vel[1] == vy and raw == 0
    # Inside the radius but moving out: pushed, velocity untouched.
    fl, pos, vel, (raw, _) = _integrate_one((0.0, FLOOR_TOP + 0.005, 0.0), (0.0, 0.6, 0.0))
    assert pos[1] == pytest.approx(FLOOR_TOP + rad, abs=1e-12) and vel[1] == 0.6 and raw == 1


def test_contacts_recorded_only_below_the_contact_depth(oracle):
    """contactDepth is 1.5 radius but the record sits inside the dw < radius branch:
    a particle between radius and contactDepth is neither pushed nor recorded."""
    rad = 0.025
    for gap, pushed in ((0.5 * rad, True), (0.99 * rad, True), (1.01 * rad, False), (1.4 * rad, False)):
        fl, pos, vel, (raw, _) = _integrate_one((0.1, FLOOR_TOP + gap, 0.1), (0.0, 0.0, 0.0))
        assert (raw == 1) == pushed, gap
        assert (pos[1] != FLOOR_TOP + gap) == pushed


def test_walls_clamp_and_zero_exactly_the_clamped_components(oracle):
    hi, lo = 1.0 - 0.025, -1.0 + 0.025
    fl, pos, vel, _ = _integrate_one((hi - 0.001, 0.5, lo + 0.001), (3.0, 0.7, -2.0))
    assert pos[0] == hi and pos[2] == lo and vel[0] == 0 and vel[2] == 0
    assert vel[1] == 0.7 and pos[1] == pytest.approx(0.5 + 0.7 / 120.0)
    fl, pos, vel, _ = _integrate_one((0.0, 0.5, 0.0), (3.0, 0.7, -2.0))
    assert np.array_equal(vel, [3.0, 0.7, -2.0])


# ---- normals -----------------------------------------------------------------------
def test_analytic_normals_match_exact_gradients(oracle):
    rng = np.random.default_rng(7)
    # sphere of radius 1 at c: the gradient is (p - c) / |p - c|
    c = np.array([0.5, 0.25, -0.75])
    sph = O.Collider().add_analytic([make_brush({"kind": "sphere", "size": 1.0, "pos": tuple(c)})])
    d = rng.normal(size=(500, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    p = c + d * rng.uniform(0.6, 1.4, size=(500, 1))
    dist, nrm = sph.eval(p, precision="f64")
    assert np.allclose(dist, np.linalg.norm(p - c, axis=1) - 1.0, atol=1e-6)
    assert np.allclose(nrm, d, atol=1e-6)
    # box faces: the exact normal of the nearest face
    box = O.Collider().add_analytic([make_brush({"kind": "box", "size": (1.0, 2.0, 3.0)})])
    for axis, half in ((0, 1.0), (1, 2.0), (2, 3.0)):
        for sgn in (1.0, -1.0):
            q = rng.uniform(-0.3, 0.3, size=(50, 3))
            q[:, axis] = sgn * (half + rng.uniform(-0.2, 0.2, size=50))
            _, n = box.eval(q, precision="f64")
            want = np.zeros(3)
            want[axis] = sgn
            assert np.allclose(n, want, atol=1e-12)
    # far inside a roundBox the central difference degenerates only at the centre
    _, n = floor_collider().eval(np.array([[0.3, FLOOR_TOP - 0.1, 0.2]]), precision="f64")
    assert np.allclose(n, [[0.0, 1.0, 0.0]], atol=1e-12)


@pytest.mark.parametrize("slope", [(0.0, 1.0, 0.0), (0.5, 1.0, 0.0), (-0.25, 0.5, 1.0)])
def test_volume_normal_of_a_baked_plane_is_its_gradient(oracle, slope):
    """A plane whose texels are exact in f16 (values in steps of 1/128 of the band
    with voxel 0.25 and band 16 voxels) is reproduced exactly by trilinear
    filtering, so the central-difference normal of volume.ts:299-313 is the plane's
    unit normal to rounding."""
    res, ws, band = 32, 8.0, 16.0
    origin = np.array([-4.0, -4.0, -4.0])
    voxel = ws / res
    g = np.asarray(slope, dtype=np.float64)
    idx = np.arange(res)
    z, y, x = np.meshgrid(idx, idx, idx, indexing="ij")
    centre = np.stack([x, y, z], -1) * voxel + origin + voxel / 2
    f = centre @ g  # plane through the origin: value in world units
    tex = np.clip(f / (band * voxel), -1.0, 1.0).astype(np.float16)
    assert np.array_equal(tex.astype(np.float64) * band * voxel, np.clip(f, -band * voxel, band * voxel))
    col = O.Collider().add_volume(tex.view(np.uint16), res, tuple(origin), ws, band)
    rng = np.random.default_rng(11)
    p = rng.uniform(-2.0, 2.0, size=(400, 3))
    p -= np.outer((p @ g) / (g @ g), g) * rng.uniform(0.0, 0.9, size=(400, 1))  # near the plane
    keep = np.abs(p @ g) < 1.0  # every texel the samples touch stays inside the band
    dist, nrm = col.eval(p[keep], precision="f64")
    assert keep.sum() > 100
    assert np.allclose(dist, p[keep] @ g, atol=1e-12)
    assert np.allclose(nrm, g / np.linalg.norm(g), atol=1e-12)
