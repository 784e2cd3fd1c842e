"""The HIP ClaySolver (lib/src/sim/pbd.ts:78-424) against the oracle, bit for bit.

Every pass follows the reference's operation order with no contraction, the match
reductions keep its lane-strided sums and pairwise tree, and sin/cos come from one
polynomial shared in form by both sides, so positions, previous positions,
velocities, rest positions and the body records (com, restCom, rot) must be
identical after many steps: bodies with plasticity, a kinematic and a weightless
body, a pushed and kicked body, free particles and particles outside their body's
slice, against an analytic floor and against a union with a baked volume.
"""
import numpy as np
import pytest

import clayhip
import oracle as O
from clayhip import scenes
from clayhip.brush import make_brush
from gpu_helpers import bits_equal

pytestmark = pytest.mark.gpu


def ball(n: int, r: float, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    p = rng.normal(size=(n, 3))
    p *= (r * rng.uniform(0, 1, size=(n, 1)) ** (1 / 3)) / np.linalg.norm(p, axis=1, keepdims=True)
    return p.astype(np.float32)


def setup():
    b0 = ball(800, 0.45, 1) + np.array([0.0, 0.2, 0.0], np.float32)
    b1 = ball(200, 0.3, 2) + np.array([1.5, -0.1, 0.0], np.float32)   # kinematic
    b2 = ball(300, 0.35, 3) + np.array([-1.4, 0.5, 0.6], np.float32)  # weightless, pushed
    free = ball(100, 0.5, 4) + np.array([0.0, 1.5, -1.5], np.float32)
    pos = np.concatenate([b0, b1, b2, free]).astype(np.float32)
    rest = pos - pos.mean(0) * np.float32(0.1)
    owner = np.concatenate([np.zeros(800), np.ones(200), np.full(300, 2), np.full(60, 0xFFFFFFFF),
                            np.zeros(40)]).astype(np.uint32)  # the last 40 claim body 0 outside its slice
    bodies = [dict(first=0, count=800, stiffness=0.55, plasticity=0.02),
              dict(first=800, count=200, stiffness=0.5, flags=1),
              dict(first=1000, count=300, stiffness=0.8, flags=2)]
    return pos, rest, owner, bodies


def run(root, world, col, steps: int):
    pos, rest, owner, bodies = setup()
    cap = 1500
    opts = dict(radius=0.06, friction=0.35, damping=0.997, maxSpeed=26.0, rotationIterations=2)
    g = clayhip.ClaySolver(root, world, opts, capacity=cap, maxBodies=5)
    for b in bodies:
        g.addBody(b["first"], b["count"], b.get("stiffness", 0.5), b.get("plasticity", 0.0), b.get("flags", 0))
    g.set_particles(pos, rest, owner)
    o = O.Solver(cap, 5, radius=0.06, friction=0.35, damping=0.997, max_speed=26.0, collider=col, precision="f32")
    o.set_bodies(bodies)
    o.set_particles(pos, rest, owner)
    for step in range(steps):
        if step == 2:
            g.setForce(2, (0.0, 3.0, 1.0), (0.5, 1.5, 0.0))
            o.set_force(2, (0.0, 3.0, 1.0), (0.5, 1.5, 0.0))
        g.step()
        o.step()
        if step in (0, 3, steps - 1):
            for name, w in (("pos", 0), ("prev", 1), ("vel", 2), ("velPrev", 3), ("rest", 4)):
                got, want = g.get(name), o.get(w)
                assert bits_equal(got, want), f"step {step} {name}: {(got != want).any(axis=1).sum()} particles differ"
            gb = g.bodies()[:3]
            assert bits_equal(gb, o.state[:3].astype(np.float32)), f"step {step} bodies: {gb} vs {o.state[:3]}"
    g.destroy()


def test_solver_on_an_analytic_floor_matches_the_oracle(root):
    brushes = [scenes.floor_brush()]
    world = clayhip.analyticField(root, brushes)
    col = O.Collider().add_analytic([make_brush(b) for b in brushes])
    run(root, world, col, 40)
    world.destroy()


def test_solver_against_a_union_with_a_volume_matches_the_oracle(root):
    brushes = scenes.ridge_world_brushes()
    vol = clayhip.SdfVolume(root, resolution=64, mipLevels=3, worldSize=8, origin=(-4, -2, -4), band=4)
    vol.bakeBrushes(brushes)
    fv = clayhip.volumeField(vol)
    fa = clayhip.analyticField(root, [scenes.floor_brush()])
    world = clayhip.unionField(fa, fv)
    col = (O.Collider().add_analytic([make_brush(scenes.floor_brush())])
           .add_volume(vol.download(0)[..., 0], 64, (-4, -2, -4), 8, 4))
    run(root, world, col, 25)
    for x in (world, fa, fv):
        x.destroy()
    vol.destroy()
