"""The tracer reads of a field on the GPU (lib/src/trace/field.ts:20-315) against the oracle.

sample and field at every mip (and above the top one) of a baked volume, an
analytic field and their union, and the entry of swept rays into a union of two
volumes, must equal the oracle's f32 mode bit for bit; bandWorld and voxelWorld
must be the f32 products the reference's shader forms.  A field with an analytic
leaf has no entry, and asking for one is an error.
"""
import numpy as np
import pytest

import clayhip
import oracle as O
from clayhip import _lib as L
from clayhip import scenes
from clayhip.brush import make_brush
from gpu_helpers import bits_equal

pytestmark = pytest.mark.gpu

SPHERE = {"kind": "sphere", "size": 0.9, "pos": (0.5, 0.6, 0.2)}


def baked(root, origin=(-4.0, -2.0, -4.0), world=8.0):
    v = clayhip.SdfVolume(root, resolution=64, mipLevels=3, worldSize=world, origin=origin, band=4)
    v.bakeBrushes([dict(b, material=float(i + 2)) for i, b in enumerate(scenes.ridge_world_brushes())])
    return v


def oracle_volume(col, vol, origin, world):
    lv = [(vol.download(m)[..., 0], vol.download(m)[..., 1]) for m in range(3)]
    return col.add_volume(lv[0][0], 64, origin, world, 4, mat16=lv[0][1], levels=lv)


def test_sample_and_field_at_every_mip(root):
    vol = baked(root)
    fv = clayhip.volumeField(vol)
    fa = clayhip.analyticField(root, [SPHERE], {"material": 7.0, "band": 3.0})
    fu = clayhip.unionField(fa, fv)
    ov = oracle_volume(O.Collider(), vol, (-4.0, -2.0, -4.0), 8.0)
    oa = O.Collider().add_analytic([make_brush(SPHERE)], material=7.0, band=3.0)
    ou = oracle_volume(O.Collider().add_analytic([make_brush(SPHERE)], material=7.0, band=3.0), vol,
                       (-4.0, -2.0, -4.0), 8.0)
    rng = np.random.default_rng(0)
    p = rng.uniform(-4.5, 4.5, size=(20000, 3)).astype(np.float32)
    p[:, 1] = rng.uniform(-2.5, 2.5, size=20000).astype(np.float32)
    for f, o, name in ((fv, ov, "volume"), (fa, oa, "analytic"), (fu, ou, "union")):
        for mip in (0, 1, 2, 3):
            for what in ("sample", "field"):
                got = getattr(f, what)(p, mip)
                want = o.query(p, mip, what, "f32")
                assert bits_equal(got, want), f"{name} {what} mip {mip}: {(got != want).any(1).sum()} points differ"
    # The union picks pairs from both children.
    s = fu.sample(p, 1)
    assert (s[:, 1] == np.clip(fa.sample(p, 1)[:, 1], -1, 1)).any() and (s == fv.sample(p, 1)).all(1).any()
    for mip in (0, 2, 5):
        vw = np.float32(8.0 / 64) * np.float32(2.0 ** min(mip, 2))
        assert fv.voxelWorld(mip) == vw and fv.bandWorld(mip) == np.float32(4.0) * np.float32(8.0 / 64) * np.float32(2.0 ** min(mip, 2))
        assert fa.bandWorld(mip) == np.float32(3.0) and fa.voxelWorld(mip) == np.float32(1e-3)
        assert fu.bandWorld(mip) == min(fa.bandWorld(mip), fv.bandWorld(mip))
        assert fu.voxelWorld(mip) == max(fa.voxelWorld(mip), fv.voxelWorld(mip))
    with pytest.raises(L.ClayError, match="no entry"):
        fu.entry(np.zeros((1, 3)), np.ones((1, 3)))
    for x in (fu, fa, fv):
        x.destroy()
    vol.destroy()


def test_entry_of_a_union_of_volumes(root):
    va = baked(root)
    vb = baked(root, origin=(3.0, -1.0, -2.0), world=4.0)
    u = clayhip.unionField(clayhip.volumeField(va), clayhip.volumeField(vb))
    ou = oracle_volume(oracle_volume(O.Collider(), va, (-4.0, -2.0, -4.0), 8.0), vb, (3.0, -1.0, -2.0), 4.0)
    rng = np.random.default_rng(1)
    n = 20000
    ro = rng.uniform(-12, 12, size=(n, 3)).astype(np.float32)
    rd = rng.normal(size=(n, 3)).astype(np.float32)
    rd[: n // 20, 0] = 0.0                      # rays parallel to a pair of faces
    tmin = rng.uniform(0, 0.5, size=n).astype(np.float32)
    radius = rng.uniform(0, 0.4, size=n).astype(np.float32)
    aperture = rng.uniform(0, 0.05, size=n).astype(np.float32)
    got = u.entry(ro, rd, tmin, radius, aperture)
    rays = np.concatenate([ro, rd, tmin[:, None], radius[:, None], aperture[:, None]], axis=1)
    want = ou.entry(rays, "f32")
    assert bits_equal(got, want), f"{(got != want).sum()} rays differ"
    assert (got == np.float32(1e9)).any() and (got < 1e9).any()
    u.destroy()
    va.destroy()
    vb.destroy()


def test_offset_and_lerp_fields_follow_device_values(root):
    """offsetField (field.ts:388-429) over lerpField (field.ts:316-386) of a morph pair,
    in a union with an analytic field, as SoftBody builds its body field
    (game/softbody.ts:209-223): sample and field at every mip, the mip-0 distance and
    normal the colliders read, and entry, equal the oracle's f32 mode bit for bit;
    then again after the blend and the offset are changed in place on the device."""
    import torch

    va = baked(root)
    vb = clayhip.SdfVolume(root, resolution=64, mipLevels=3, worldSize=8.0, origin=(-4.0, -2.0, -4.0), band=4)
    vb.bakeBrushes([{"kind": "sphere", "size": 1.4, "pos": (-0.8, -0.4, 0.5), "material": 5.0},
                    {"kind": "sphere", "size": 0.9, "pos": (1.5, 0.2, -0.6), "material": 6.0}])
    t = torch.tensor([0.35], dtype=torch.float32, device="cuda")
    off = torch.tensor([0.3, -0.2, 0.45], dtype=torch.float32, device="cuda")
    fa, fb = clayhip.volumeField(va), clayhip.volumeField(vb)
    lerp = clayhip.lerpField(fa, fb, t)
    moved = clayhip.offsetField(lerp, off)
    ana = clayhip.analyticField(root, [SPHERE], {"material": 7.0, "band": 3.0})
    u = clayhip.unionField(ana, moved)
    rng = np.random.default_rng(4)
    p = rng.uniform(-4.5, 4.5, size=(20000, 3)).astype(np.float32)
    rays = np.concatenate([rng.uniform(-12, 12, size=(5000, 3)), rng.normal(size=(5000, 3)),
                           rng.uniform(0, 0.5, size=(5000, 1)), rng.uniform(0, 0.4, size=(5000, 1)),
                           rng.uniform(0, 0.05, size=(5000, 1))], axis=1).astype(np.float32)

    def oracles(tv, ov):
        pair = lambda: (oracle_volume(O.Collider(), va, (-4.0, -2.0, -4.0), 8.0),
                        oracle_volume(O.Collider(), vb, (-4.0, -2.0, -4.0), 8.0))
        ol = O.Collider().add_lerp(*pair(), tv)
        om = O.Collider().add_lerp(*pair(), tv).offset(ov)
        ou = O.Collider().add_analytic([make_brush(SPHERE)], material=7.0, band=3.0).add_lerp(*pair(), tv)
        ou.c.leaf[1].has_offset = 1
        for a in range(3):
            ou.c.leaf[1].offset[a] = float(ov[a])
        return ol, om, ou

    def check(tv, ov):
        ol, om, ou = oracles(tv, ov)
        for f, o, name in ((lerp, ol, "lerp"), (moved, om, "offset lerp"), (u, ou, "union")):
            for mip in (0, 1, 2, 3):
                for what in ("sample", "field"):
                    got, want = getattr(f, what)(p, mip), o.query(p, mip, what, "f32")
                    assert bits_equal(got, want), f"{name} {what} mip {mip}: {(got != want).any(1).sum()} differ"
            d, n = f.evaluate(p)
            od, on = o.eval(p)
            assert bits_equal(d, od) and bits_equal(n, on), f"{name} mip-0 distance or normal differs"
        got = moved.entry(rays[:, :3], rays[:, 3:6], rays[:, 6], rays[:, 7], rays[:, 8])
        assert bits_equal(got, om.entry(rays, "f32")), "offset lerp entry differs"

    check(np.float32(0.35), np.array([0.3, -0.2, 0.45], np.float32))
    t.fill_(0.8)
    off.copy_(torch.tensor([-0.6, 0.1, 0.2]))
    torch.cuda.synchronize()
    check(np.float32(0.8), np.array([-0.6, 0.1, 0.2], np.float32))
    with pytest.raises(L.ClayError, match="morph pair"):
        clayhip.lerpField(u, fb, t)
    with pytest.raises(L.ClayError, match="already offset"):
        clayhip.offsetField(moved, off)
    for x in (u, ana, moved, lerp, fa, fb):
        x.destroy()
    va.destroy()
    vb.destroy()
