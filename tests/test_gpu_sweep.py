"""The tracer sweep on the GPU (lib/src/trace/march.ts:71-218) against the oracle.

* makeTracer(field, options).sweep of rays, sphere sweeps and cones through a baked
  volume, an analytic field, their union (no entry: marched from tMin), a union of
  two volumes (started at entry) and a moved morph pair, under the default and
  changed options: t, hit, steps and dist equal the oracle's f32 mode bit for bit.
* The scenes and rays of lib/test/gpu/march.gpu.ts:36-260, baked on the GPU,
  meet that file's expectations (hits, distances, step budgets) and equal the
  oracle's sweep over the same texels.
* The device entry point writes what the host one returns.
"""
import numpy as np
import pytest

import clayhip
import oracle as O
from clayhip import scenes
from clayhip.brush import make_brush
from gpu_helpers import bits_equal

pytestmark = pytest.mark.gpu

SPHERE = {"kind": "sphere", "size": 0.9, "pos": (0.5, 0.6, 0.2)}
OPTIONS = [None, {"maxSteps": 48, "minMip": 1}, {"minStep": 0.5, "hitEps": 1.0}, {"minMip": 2, "maxSteps": 300}]


def baked(root, origin=(-4.0, -2.0, -4.0), world=8.0):
    v = clayhip.SdfVolume(root, resolution=64, mipLevels=3, worldSize=world, origin=origin, band=4)
    v.bakeBrushes([dict(b, material=float(i + 2)) for i, b in enumerate(scenes.ridge_world_brushes())])
    return v


def oracle_volume(col, vol, origin, world, res=64, band=4.0):
    lv = [(vol.download(m)[..., 0], vol.download(m)[..., 1]) for m in range(vol.mipLevels)]
    return col.add_volume(lv[0][0], res, origin, world, band, mat16=lv[0][1], levels=lv)


def random_rays(seed, n=6000, span=10.0):
    rng = np.random.default_rng(seed)
    ro = rng.uniform(-span, span, size=(n, 3)).astype(np.float32)
    ro[: n // 3, 1] = rng.uniform(0.5, 3.0, size=n // 3)  # from above the ridges, looking around
    rd = rng.normal(size=(n, 3)).astype(np.float32)
    rd /= np.sqrt((rd * rd).sum(1, dtype=np.float32))[:, None]
    rd[n // 3: n // 2] = (-ro[n // 3: n // 2] / np.linalg.norm(ro[n // 3: n // 2], axis=1)[:, None]).astype(np.float32)
    tmin = rng.uniform(0, 0.5, size=n).astype(np.float32)
    tmax = rng.uniform(5, 30, size=n).astype(np.float32)
    radius = np.where(rng.random(n) < 0.3, rng.uniform(0, 0.4, size=n), 0).astype(np.float32)
    aperture = np.where(rng.random(n) < 0.3, rng.uniform(0, 0.05, size=n), 0).astype(np.float32)
    return ro, rd, tmin, tmax, radius, aperture


def check(f, o, rays, options, name):
    ro, rd, tmin, tmax, radius, aperture = rays
    got = f.sweep(ro, rd, tmin, tmax, radius, aperture, options)
    r10 = np.concatenate([ro, rd, tmin[:, None], tmax[:, None], radius[:, None], aperture[:, None]], axis=1)
    opt = options or {}
    want = o.sweep(r10, max_steps=opt.get("maxSteps", 192), min_step=opt.get("minStep", 0.25),
                   hit_eps=opt.get("hitEps", 0.5), min_mip=opt.get("minMip", 0))
    for k in ("t", "dist"):
        assert bits_equal(got[k], want[k]), f"{name} {options}: {k} differs on {(got[k] != want[k]).sum()} rays"
    for k in ("hit", "steps"):
        assert np.array_equal(got[k], want[k]), f"{name} {options}: {k} differs on {(got[k] != want[k]).sum()} rays"
    return got


def test_sweep_matches_the_oracle(root):
    import torch

    va = baked(root)
    vb = baked(root, origin=(3.0, -1.0, -2.0), world=4.0)
    vc = clayhip.SdfVolume(root, resolution=64, mipLevels=3, worldSize=8.0, origin=(-4.0, -2.0, -4.0), band=4)
    vc.bakeBrushes([{"kind": "sphere", "size": 1.4, "pos": (-0.8, -0.4, 0.5), "material": 5.0}])
    fa = clayhip.volumeField(va)
    ana = clayhip.analyticField(root, [SPHERE], {"material": 7.0, "band": 3.0})
    mixed = clayhip.unionField(ana, clayhip.volumeField(va))
    two = clayhip.unionField(clayhip.volumeField(va), clayhip.volumeField(vb))
    t = torch.tensor([0.35], dtype=torch.float32, device="cuda")
    off = torch.tensor([0.3, -0.2, 0.45], dtype=torch.float32, device="cuda")
    moved = clayhip.offsetField(clayhip.lerpField(clayhip.volumeField(va), clayhip.volumeField(vc), t), off)

    oa = oracle_volume(O.Collider(), va, (-4.0, -2.0, -4.0), 8.0)
    oana = O.Collider().add_analytic([make_brush(SPHERE)], material=7.0, band=3.0)
    omixed = oracle_volume(O.Collider().add_analytic([make_brush(SPHERE)], material=7.0, band=3.0), va,
                           (-4.0, -2.0, -4.0), 8.0)
    otwo = oracle_volume(oracle_volume(O.Collider(), va, (-4.0, -2.0, -4.0), 8.0), vb, (3.0, -1.0, -2.0), 4.0)
    omoved = O.Collider().add_lerp(oracle_volume(O.Collider(), va, (-4.0, -2.0, -4.0), 8.0),
                                   oracle_volume(O.Collider(), vc, (-4.0, -2.0, -4.0), 8.0), 0.35).offset(
                                       (0.3, -0.2, 0.45))
    cases = ((fa, oa, "volume"), (ana, oana, "analytic"), (mixed, omixed, "analytic + volume"),
             (two, otwo, "two volumes"), (moved, omoved, "moved morph pair"))
    for seed, options in enumerate(OPTIONS):
        rays = random_rays(seed)
        for f, o, name in cases:
            got = check(f, o, rays, options, name)
            if name != "analytic":
                assert got["hit"].any() and not got["hit"].all(), name
    for x in (fa, ana, mixed, two, moved):
        x.destroy()
    for v in (va, vb, vc):
        v.destroy()


def unit(v):
    v = np.asarray(v, np.float32)
    return v / np.float32(np.sqrt(np.dot(v, v)))


@pytest.fixture(scope="module")
def reference_scene(root):
    # march.gpu.ts:52-85: a sphere of radius 3 at 128^3 over 16 units, five mips; a
    # pebble of radius 1 at (5, 0, 0) in a 32^3 (two-mip) volume over 4 units.
    big = clayhip.SdfVolume(root, resolution=128, mipLevels=5, worldSize=16.0, origin=(-8.0, -8.0, -8.0), band=4)
    big.bakeBrushes([{"kind": "sphere", "size": 3.0}])
    peb = clayhip.SdfVolume(root, resolution=32, worldSize=4.0, origin=(3.0, -2.0, -2.0), band=4)
    peb.bakeBrushes([{"kind": "sphere", "size": 1.0, "pos": (5.0, 0.0, 0.0)}])
    yield big, peb
    big.destroy()
    peb.destroy()


def test_reference_tracer_scenes(reference_scene):
    big, peb = reference_scene
    assert peb.mipLevels == 2
    f = clayhip.volumeField(big)
    o = oracle_volume(O.Collider(), big, (-8.0, -8.0, -8.0), 16.0, res=128)

    def run(ro, rd, tmax, radius=0.0, field=f, orc=o):
        ro = np.asarray(ro, np.float32).reshape(-1, 3)
        rd = np.stack([unit(d) for d in np.asarray(rd, np.float32).reshape(-1, 3)])
        n = max(len(ro), len(rd))
        rays = (np.broadcast_to(ro, (n, 3)).copy(), np.broadcast_to(rd, (n, 3)).copy(), np.zeros(n, np.float32),
                np.full(n, tmax, np.float32), np.broadcast_to(np.float32(radius), (n,)).astype(np.float32),
                np.zeros(n, np.float32))
        return check(field, orc, rays, None, "reference scene")

    h = run([[0, 0, -14], [-12, 6, -12]], [[0, 0, 1], [12, -6, 12]], 60.0)  # march.gpu.ts:133-155
    assert (h["hit"] == 1).all() and abs(h["t"][0] - 11.0) < 0.2 and (h["steps"] < 20).all(), h

    near = 3.05  # march.gpu.ts:157-185
    h = run([[near, 0, 0]] * 3, [[1, 0, 0], [1, 2, 0], [-1, 0, 0]], 20.0)
    assert h["hit"][0] == 0 and h["hit"][1] == 0, h
    assert h["hit"][2] == 1 and abs(h["t"][2] - 0.05) < 0.07, h
    h = run([3.5, 0, 0], [1, 0, 0], 20.0, radius=0.75)
    assert h["hit"][0] == 1 and h["t"][0] < 0.05, h

    h = run([-9, 8.4, 6], [20, -1, 0], 40.0)  # march.gpu.ts:187-212
    assert h["hit"][0] == 0 and h["t"][0] >= 40 - 1e-3 and h["steps"][0] < 192, h

    u = clayhip.unionField(clayhip.volumeField(big), clayhip.volumeField(peb))  # march.gpu.ts:214-260
    assert u.maxMip == 4
    ou = oracle_volume(oracle_volume(O.Collider(), big, (-8.0, -8.0, -8.0), 16.0, res=128), peb,
                       (3.0, -2.0, -2.0), 4.0, res=32)
    h = run([[12, 0, 0], [12, 1.6, 0], [12, 5, 0]], [-1, 0, 0], 30.0, field=u, orc=ou)
    assert h["hit"][0] == 1 and abs(h["t"][0] - 6.0) < 0.25, h
    assert h["hit"][1] == 1 and abs(h["t"][1] - (12 - np.sqrt(9 - 2.56))) < 0.25, h
    assert h["hit"][2] == 0 and h["steps"][2] < 25, h
    u.destroy()
    f.destroy()


def test_device_entry_point_matches_the_host_one(reference_scene):
    import torch

    big, _ = reference_scene
    f = clayhip.volumeField(big)
    ro, rd, tmin, tmax, radius, aperture = random_rays(7, n=4096, span=12.0)
    r10 = np.concatenate([ro, rd, tmin[:, None], tmax[:, None], radius[:, None], aperture[:, None]], axis=1)
    want = f.sweep(ro, rd, tmin, tmax, radius, aperture, {"maxSteps": 96})
    rays = torch.from_numpy(r10).cuda()
    hits = torch.zeros((len(r10), 4), dtype=torch.int32, device="cuda")
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        f.sweep_device(rays.data_ptr(), len(r10), hits.data_ptr(), stream.cuda_stream, {"maxSteps": 96})
    stream.synchronize()
    h = hits.cpu().numpy()
    assert bits_equal(h[:, 0].view(np.float32), want["t"]) and bits_equal(h[:, 3].view(np.float32), want["dist"])
    assert np.array_equal(h[:, 1].view(np.uint32), want["hit"]) and np.array_equal(h[:, 2].view(np.uint32), want["steps"])
    f.destroy()
