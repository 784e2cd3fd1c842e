"""The oracle's tracer sweep (lib/src/trace/march.ts:71-218), pinned on the CPU.

* On an analytic sphere, the properties lib/test/march.test.mjs:35-124 checks of the
  same loop: a ray stops on the surface it should hit, a miss reports no hit and
  runs to tMax, no ray, sphere sweep or cone passes the tangency point
  (t + radius + aperture t = 8 on the axis), and the snap leaves the swept shape
  clear of a field whose nearest feature changes along the ray.
* Saturated readings are never hits: an empty baked volume (every texel +1) is
  crossed with no hit.
* The hierarchy, on the scenes and rays of lib/test/gpu/march.gpu.ts:36-260 (a
  sphere of radius 3 baked at 128^3 over 16 units with five mips, and a 32^3
  two-mip pebble), with that file's expected hits, distances and step budgets.
* The level bookkeeping against a count made by hand: with a maximum of 0 steps
  nothing is sampled; the start level is the deepest even offset above minMip.
"""
import numpy as np
import pytest

import oracle as O
from clayhip.brush import make_brush


def sphere_collider(center=(0.0, 0.0, 10.0), r=2.0, extra=()):
    brushes = [make_brush({"kind": "sphere", "size": r, "pos": center, "smooth": 0.0})]
    brushes += [make_brush(b) for b in extra]
    return O.Collider().add_analytic(brushes)


def rays(ro, rd, tmin=0.0, tmax=100.0, radius=0.0, aperture=0.0):
    ro = np.atleast_2d(np.asarray(ro, np.float64))
    rd = np.atleast_2d(np.asarray(rd, np.float64))
    n = max(ro.shape[0], rd.shape[0])
    out = np.zeros((n, 10))
    out[:, 0:3] = ro
    out[:, 3:6] = rd
    out[:, 6] = tmin
    out[:, 7] = tmax
    out[:, 8] = radius
    out[:, 9] = aperture
    return out


def unit(v):
    v = np.asarray(v, np.float32)
    return v / np.float32(np.sqrt(np.dot(v, v)))


def test_ray_stops_on_the_sphere_and_a_miss_runs_to_tmax():
    col = sphere_collider()
    for prec in ("f32", "f64"):
        h = col.sweep(rays([0, 0, 0], [0, 0, 1]), precision=prec)
        assert h["hit"][0] == 1
        # hitEps is half a voxel of the analytic field (eps 1e-3): the snap lands
        # within it of the front face at t = 8.
        assert abs(h["t"][0] - 8.0) < 1e-3, h["t"][0]
        m = col.sweep(rays([0, 0, 0], [0, 1, 0], tmax=50.0), precision=prec)
        assert m["hit"][0] == 0 and m["t"][0] == 50.0


@pytest.mark.parametrize("aperture", [0.0, 0.01, 0.05, 0.2])
@pytest.mark.parametrize("radius", [0.0, 0.25, 1.0])
def test_no_sweep_passes_the_tangency_point(radius, aperture):
    col = sphere_collider()
    h = col.sweep(rays([0, 0, 0], [0, 0, 1], radius=radius, aperture=aperture), precision="f64")
    tangency = (8.0 - radius) / (1.0 + aperture)
    assert h["hit"][0] == 1
    assert h["t"][0] <= tangency + 1e-9
    # ... and stops within the hit threshold of it (hitEps = 5e-4 on this field).
    assert tangency - h["t"][0] < 2e-3


@pytest.mark.parametrize("aperture", [0.0, 0.02, 0.1])
@pytest.mark.parametrize("radius", [0.0, 0.5])
def test_snap_keeps_the_swept_shape_clear_when_the_nearest_feature_changes(radius, aperture):
    # A sphere in front, the half-space y > 12 behind it (a plane brush turned over).
    col = sphere_collider(center=(0.0, 10.0, 0.0),
                          extra=[{"kind": "plane", "pos": (0, 12, 0), "euler": (np.pi, 0, 0), "smooth": 0.0}])
    for ro, rd in (([0, 0, 0], [0, 1, 0]), ([2.5, 0, 0], [0, 1, 0]), ([0.1, 0, 1.9], list(unit([0.02, 1, 0.01])))):
        h = col.sweep(rays(ro, rd, radius=radius, aperture=aperture), precision="f64")
        assert h["hit"][0] == 1
        p = np.asarray(ro) + np.asarray(rd) * h["t"][0]
        clearance = col.query(p[None], 0, "sample", "f64")[0, 0]
        assert clearance >= radius + aperture * h["t"][0] - 1e-9, (ro, clearance)
        assert clearance < radius + aperture * h["t"][0] + 1e-3, (ro, clearance)


def test_saturated_readings_are_never_hits():
    # An empty bake: every texel at +1 (saturated); the box itself reads saturated
    # from outside.  Rays through it, grazing it and starting inside run to tMax.
    lv = O.bake_brushes(32, 2, 4.0, (-2.0, -2.0, -2.0), 4.0, [])
    levels = [(x[..., 0], x[..., 1]) for x in lv]
    col = O.Collider().add_volume(levels[0][0], 32, (-2.0, -2.0, -2.0), 4.0, 4.0, mat16=levels[0][1], levels=levels)
    r = rays([[-5, 0, 0], [-5, 2.0001, 0], [0, 0, 0], [0.5, -1, 0.3]],
             [[1, 0, 0], [1, 0, 0], list(unit([1, 2, 3])), list(unit([-1, 0.2, 0.1]))], tmax=12.0)
    h = col.sweep(r)
    assert (h["hit"] == 0).all() and (h["t"] == 12.0).all()
    assert (h["steps"] < 60).all(), h["steps"]


def baked_sphere(res=128, mips=5, world=16.0, origin=(-8.0, -8.0, -8.0), r=3.0, center=(0.0, 0.0, 0.0),
                 col=None, band=4.0):
    b = make_brush({"kind": "sphere", "size": r, "pos": center})
    lv = O.bake_brushes(res, mips, world, origin, band, [b])
    levels = [(x[..., 0], x[..., 1]) for x in lv]
    col = col or O.Collider()
    return col.add_volume(levels[0][0], res, origin, world, band, mat16=levels[0][1], levels=levels)


@pytest.fixture(scope="module")
def scene():
    return baked_sphere()


def test_flying_in_from_outside_costs_a_handful_of_steps(scene):
    # march.gpu.ts:133-155
    h = scene.sweep(rays([[0, 0, -14], [-12, 6, -12]], [[0, 0, 1], list(unit([12, -6, 12]))], tmax=60.0))
    assert (h["hit"] == 1).all()
    assert abs(h["t"][0] - 11.0) < 0.2, h["t"]
    assert (h["steps"] < 20).all(), h["steps"]


def test_a_ray_inside_hit_eps_of_the_surface_may_leave_it(scene):
    # march.gpu.ts:157-185
    near = 3.05
    r = rays([[near, 0, 0], [near, 0, 0], [near, 0, 0], [3.5, 0, 0]],
             [[1, 0, 0], list(unit([1, 2, 0])), [-1, 0, 0], [1, 0, 0]], tmax=20.0)
    r[3, 8] = 0.75
    h = scene.sweep(r)
    assert h["hit"][0] == 0 and h["hit"][1] == 0
    assert h["hit"][2] == 1 and abs(h["t"][2] - 0.05) < 0.07, h["t"][2]
    assert h["hit"][3] == 1 and h["t"][3] < 0.05, h["t"][3]


def test_a_grazing_ray_crosses_the_volume(scene):
    # march.gpu.ts:187-212
    h = scene.sweep(rays([-9, 8.4, 6], list(unit([20, -1, 0])), tmax=40.0))
    assert h["hit"][0] == 0
    assert h["t"][0] >= 40 - 1e-3
    assert h["steps"][0] < 192


def test_a_shallow_volume_joins_a_deep_one(scene):
    # march.gpu.ts:214-260: the union keeps the deeper hierarchy (maxMip 4), the
    # pebble's queries clamp to its own two levels.
    u = baked_sphere(col=baked_sphere(), res=32, mips=2, world=4.0, origin=(3.0, -2.0, -2.0), r=1.0,
                     center=(5.0, 0.0, 0.0))
    h = u.sweep(rays([[12, 0, 0], [12, 1.6, 0], [12, 5, 0]], [-1, 0, 0], tmax=30.0))
    assert h["hit"][0] == 1 and abs(h["t"][0] - 6.0) < 0.25, h["t"]
    assert h["hit"][1] == 1 and abs(h["t"][1] - (12 - np.sqrt(9 - 2.56))) < 0.25, h["t"]
    assert h["hit"][2] == 0 and h["steps"][2] < 25, h["steps"]


def test_level_bookkeeping(scene):
    # No iteration: nothing sampled, t = min(tStart, tMax) with tStart the entry.
    h = scene.sweep(rays([0, 0, -14], [0, 0, 1], tmin=0.5, tmax=60.0), max_steps=0)
    assert h["steps"][0] == 0 and h["hit"][0] == 0 and h["dist"][0] == 0
    assert h["t"][0] == np.float32(6.0)  # the box face at z = -8, 6 units out
    # One iteration from the entry point samples the start level: maxMip 4 with
    # minMip 0 starts on 4; minMip 1 starts on 3 (4 - 1 is odd, so one level lower).
    p = np.array([[0.0, 0.0, -8.0]])
    for min_mip, start in ((0, 4), (1, 3), (4, 4), (9, 4)):
        h = scene.sweep(rays([0, 0, -14], [0, 0, 1], tmax=60.0), max_steps=1, min_mip=min_mip)
        assert h["steps"][0] == 1
        assert h["dist"][0] == scene.query(p, start, "sample", "f32")[0, 0]


def test_a_hop_steps_back_half_a_finer_voxel(scene):
    # march.gpu.ts's head-on ray: the first two samples (mip 4) promise empty space
    # and the ray advances; the third no longer does, so the trace drops to mip 2 and
    # backs off half a mip-2 voxel (0.5 * 16/128 * 4 = 0.25) before the fourth
    # sample, which reads mip 2 at the backed-off point.
    r = rays([0, 0, -14], [0, 0, 1], tmax=60.0)
    h = [scene.sweep(r, max_steps=k) for k in (1, 2, 3, 4)]
    t = [x["t"][0] for x in h]
    assert 6.0 < t[0] < t[1] and not any(x["hit"][0] for x in h)
    assert t[2] == np.float32(t[1] - np.float32(0.25))
    p = np.array([[0.0, 0.0, -14.0 + float(t[2])]], np.float32)
    assert h[3]["dist"][0] == scene.query(p, 2, "sample", "f32")[0, 0]
