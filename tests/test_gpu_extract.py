"""The HIP SurfaceExtractor (lib/src/sim/extract.ts:85-484) against the oracle, bit for bit.

Both take particle slots in ascending cell index and quads in ascending lattice
point then edge axis, and evaluate every corner, crossing, snap and material in the
reference's operation order, so counts, draw arguments, the index list, every
particle plane (pos, prev, vel, velPrev, rest, restNormal, material), the owners and
the body record must be identical.  Covered: an analytic sphere with a material; a
union of a floor and a baked volume; a capacity that cuts the extraction; a motion
record with preserveMotion; and a ClaySolver step after the extraction.
"""
import numpy as np
import pytest

import clayhip
import oracle as O
from clayhip import scenes
from clayhip.brush import make_brush
from gpu_helpers import bits_equal

pytestmark = pytest.mark.gpu

SPHERE = {"kind": "sphere", "size": 0.7, "pos": (0.1, 0.05, -0.02)}


def compare(g, o, ge, oe, nbodies):
    counts, draw, idx = ge.read()
    assert np.array_equal(counts, oe.counts), f"counts {counts} vs {oe.counts}"
    assert np.array_equal(draw, oe.draw), f"draw args {draw} vs {oe.draw}"
    assert np.array_equal(idx, oe.indices), "quad indices differ"
    for name, w in (("pos", 0), ("prev", 1), ("vel", 2), ("velPrev", 3), ("rest", 4)):
        got, want = g.get(name), o.get(w)
        assert bits_equal(got, want), f"{name}: {(got != want).any(axis=1).sum()} particles differ"
    assert bits_equal(g.get("restNormal"), o.aux[:3].T), "restNormal differs"
    assert bits_equal(g.get("material"), o.aux[3]), "material differs"
    assert np.array_equal(g.owners(), o.owner), "owners differ"
    assert np.array_equal(g.slices()[:nbodies], o.body3[:nbodies]), f"slices {g.slices()[:nbodies]} vs {o.body3[:nbodies]}"
    assert bits_equal(g.bodies()[:nbodies], o.state[:nbodies]), "body records differ"


def make_pair(root, world, col, cap=4096, bodies=2):
    g = clayhip.ClaySolver(root, world, dict(radius=0.06), capacity=cap, maxBodies=bodies)
    o = O.Solver(cap, bodies, radius=0.06, collider=col, precision="f32")
    return g, o


def test_sphere_extraction_matches_the_oracle(root):
    field = clayhip.analyticField(root, [SPHERE], {"material": 3.0})
    col = O.Collider().add_analytic([make_brush(SPHERE)], material=3.0)
    g, o = make_pair(root, field, col)
    # Body 0: the whole surface at 24^3, with an emission velocity.
    assert g.addBody(stiffness=0.55) == o.add_body(stiffness=0.55) == 0
    ge = clayhip.SurfaceExtractor(g, field, dict(resolution=24, body=0, base=0, capacity=3000))
    oe = O.Extractor(o, col, resolution=24, body=0, base=0, capacity=3000)
    ge.setRegion((-1.0, -1.0, -1.0), 2.0, (0.25, -0.5, 1.0))
    oe.set_region((-1.0, -1.0, -1.0), 2.0, (0.25, -0.5, 1.0))
    ge.extract()
    oe.extract()
    compare(g, o, ge, oe, 1)
    # Body 1: a 40^3 lattice cut to 250 particles in slots [3000, 3250).
    assert g.addBody() == o.add_body() == 1
    ge2 = clayhip.SurfaceExtractor(g, field, dict(resolution=40, body=1, base=3000, capacity=250))
    oe2 = O.Extractor(o, col, resolution=40, body=1, base=3000, capacity=250)
    ge2.setRegion((-0.9, -0.8, -0.95), 1.9)
    oe2.set_region((-0.9, -0.8, -0.95), 1.9)
    ge2.extract()
    oe2.extract()
    assert oe2.counts[0] > 250
    compare(g, o, ge2, oe2, 2)
    for e in (ge, ge2):
        e.destroy()
    g.destroy()
    field.destroy()


def test_union_with_a_volume_and_a_solver_step(root):
    vol = clayhip.SdfVolume(root, resolution=64, mipLevels=3, worldSize=8, origin=(-4, -2, -4), band=4)
    ridge = [dict(b, material=4.0) for b in scenes.ridge_world_brushes()]
    vol.bakeBrushes(ridge)
    fv = clayhip.volumeField(vol)
    fa = clayhip.analyticField(root, [SPHERE], {"material": 1.0})
    world = clayhip.unionField(fa, fv)
    tex = vol.download(0)
    col = (O.Collider().add_analytic([make_brush(SPHERE)], material=1.0)
           .add_volume(tex[..., 0], 64, (-4, -2, -4), 8, 4, mat16=tex[..., 1]))
    g, o = make_pair(root, world, col, bodies=1)
    g.addBody(stiffness=0.6, plasticity=0.01)
    o.add_body(stiffness=0.6, plasticity=0.01)
    ge = clayhip.SurfaceExtractor(g, world, dict(resolution=32, body=0, base=0, capacity=4000))
    oe = O.Extractor(o, col, resolution=32, body=0, base=0, capacity=4000)
    ge.setRegion((-1.6, -0.4, -1.6), 3.2)
    oe.set_region((-1.6, -0.4, -1.6), 3.2)
    ge.extract()
    oe.extract()
    compare(g, o, ge, oe, 1)
    assert (g.get("material")[: int(oe.counts[0])] > 0).all()
    for _ in range(3):
        g.step()
        o.step()
    compare(g, o, ge, oe, 1)
    ge.destroy()
    g.destroy()
    for x in (world, fa, fv):
        x.destroy()
    vol.destroy()


def test_motion_record_and_preserve_motion(root):
    field = clayhip.analyticField(root, [SPHERE])
    col = O.Collider().add_analytic([make_brush(SPHERE)])
    g, o = make_pair(root, field, col, bodies=1)
    g.addBody()
    o.add_body()
    ge = clayhip.SurfaceExtractor(g, field, dict(resolution=20, body=0, base=0, capacity=3000, motion=True))
    oe = O.Extractor(o, col, resolution=20, body=0, base=0, capacity=3000, motion=True)
    rec = np.array([0.05, -0.1, 0.02, 0.0, 0.5, 0.0, -0.25, 0.0], np.float32)
    ge.motion(rec)
    oe.motion[:] = rec
    ge.setRegion((-1.0, -1.0, -1.0), 2.0)
    oe.set_region((-1.0, -1.0, -1.0), 2.0)
    ge.extract()
    oe.extract()
    compare(g, o, ge, oe, 1)
    for _ in range(4):      # let the body fall, then re-extract keeping its motion
        g.step()
        o.step()
    ge.setRegion((-1.0, -1.0, -1.0), 2.0, (0.0, 0.0, 0.0), True)
    oe.set_region((-1.0, -1.0, -1.0), 2.0, preserve_motion=True)
    ge.extract()
    oe.extract()
    assert bits_equal(ge.motion(), oe.motion), f"motion {ge.motion()} vs {oe.motion}"
    compare(g, o, ge, oe, 1)
    ge.destroy()
    g.destroy()
    field.destroy()
