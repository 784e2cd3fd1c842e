"""The oracle's ClaySolver (lib/src/sim/pbd.ts:78-424), pinned on the CPU.

* sin/cos (quatFromAxisAngle) against libm;
* extractRotation (math/gpu.ts:55-78) recovering a known rotation;
* predict + finalize against the BDF2 recurrence written out for one particle
  of a rigid body in free fall (match and apply must not disturb a translation);
* invariants: a body dropped on a floor ends at rest one radius above it,
  kinematic and weightless bodies, ownership (free particles and particles
  outside their body's slice are never touched), full plasticity.
"""
import math

import numpy as np
import pytest

import oracle as O
from clayhip.brush import make_brush
from clayhip import scenes


def ball(n: int, r: float, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    p = rng.normal(size=(n, 3))
    p *= (r * rng.uniform(0, 1, size=(n, 1)) ** (1 / 3)) / np.linalg.norm(p, axis=1, keepdims=True)
    return p.astype(np.float32)


def quat_to_mat(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_sin_cos_against_libm(precision):
    xs = np.concatenate([np.linspace(-12.0, 12.0, 20001), np.array([0.0, 1e-8, -1e-8, math.pi / 4, math.pi])])
    tol = 2.5e-7 if precision == "f32" else 3e-9  # f64: the same 9-digit coefficients
    for x in xs:
        s, c = O.sin_cos(float(x), precision)
        xr = float(np.float32(x)) if precision == "f32" else float(x)
        assert abs(s - math.sin(xr)) <= tol * max(1.0, abs(xr) / 4), x
        assert abs(c - math.cos(xr)) <= tol * max(1.0, abs(xr) / 4), x


def test_extract_rotation_recovers_a_known_rotation():
    """A rigid body turned by R0 about its centroid: the match pass must find R0."""
    rest = ball(600, 0.8, 1)
    axis = np.array([0.3, -0.5, 0.8])
    axis /= np.linalg.norm(axis)
    ang = 0.7
    q0 = np.concatenate([axis * math.sin(ang / 2), [math.cos(ang / 2)]])
    pos = (rest.astype(np.float64) - rest.mean(0)) @ quat_to_mat(q0).T + rest.mean(0) + np.array([1.0, 2.0, -0.5])
    s = O.Solver(600, 2, gravity=(0, 0, 0), substeps=1, rotation_iterations=30, precision="f64")
    s.set_bodies([dict(first=0, count=600, stiffness=1.0, flags=2)])
    s.set_particles(pos, rest, np.zeros(600, np.uint32))
    s.step()
    q = s.state[0, 6:10]
    if q[3] < 0:
        q = -q
    assert np.allclose(q, q0, atol=1e-9)
    assert np.allclose(s.state[0, 0:3], pos.mean(0), atol=1e-12)
    assert np.allclose(s.state[0, 3:6], rest.astype(np.float64).mean(0), atol=1e-12)
    # Already rigid: shape matching leaves every particle where it was.
    assert np.allclose(s.get(0)[:600], pos, atol=1e-12)


def test_free_fall_follows_the_bdf2_recurrence():
    g, dt, sub, damp = -9.81, 1 / 60, 2, 0.995
    h = dt / sub
    rest = ball(300, 0.5, 2)
    s = O.Solver(300, 1, gravity=(0, g, 0), dt=dt, substeps=sub, damping=damp, max_speed=1e9, precision="f64")
    s.set_bodies([dict(first=0, count=300, stiffness=0.5)])
    s.set_particles(rest, rest, np.zeros(300, np.uint32))
    # One particle's y: x' = (4x - xp)/3 + (2h/3)((4v - vp)/3 + (2h/3)g); v' = (x' - x)/h * damp.
    x = xp = 0.0
    v = vp = 0.0
    for _ in range(10):
        s.step()
        for _ in range(sub):
            vpred = (4 * v - vp) / 3 + g * (2 * h / 3)
            xn = (4 * x - xp) / 3 + vpred * (2 * h / 3)
            xp, x = x, xn
            vp, v = v, (x - xp) / h * damp
    got = s.get(0)[:300] - rest
    assert np.allclose(got[:, 1], x, atol=1e-11) and np.allclose(got[:, [0, 2]], 0, atol=1e-11)
    assert np.allclose(s.get(2)[:300, 1], v, atol=1e-9)


def test_body_comes_to_rest_on_a_floor():
    """Dropped from 0.3 above a roundBox floor (top at y = -0.4), the body settles: its
    centroid stops moving and no particle ends closer to the surface than the solver
    radius.  (The contact particles keep a downward velocity: the pushout moves `prev`
    with `pos`, pbd.ts:201-219, so the predicted motion survives into the velocity and
    only the speed cap bounds it; the positions are projected out every substep.)"""
    floor = make_brush(scenes.floor_brush())
    col = O.Collider().add_analytic([floor])
    rest = ball(400, 0.4, 3)
    pos = rest + np.array([0.0, 0.4 + 0.3 - 0.4, 0.0], dtype=np.float32)
    s = O.Solver(400, 1, radius=0.05, collider=col, precision="f32")
    s.set_bodies([dict(first=0, count=400, stiffness=0.55)])
    s.set_particles(pos, rest, np.zeros(400, np.uint32))
    coms = []
    for _ in range(240):
        s.step()
        coms.append(s.state[0, 1])
    assert np.ptp(coms[-100:]) < 5e-3
    p = s.get(0)[:400]
    d = np.empty(400, np.float32)
    nrm = np.empty((400, 3), np.float32)
    O.lib().orc_collider_eval_f32(col.c, 400, np.ascontiguousarray(p).ctypes.data, d.ctypes.data, nrm.ctypes.data)
    assert d.min() > 0.05 - 5e-3


def test_kinematic_weightless_and_ownership():
    rest = ball(200, 0.3, 4)
    owner = np.zeros(600, np.uint32)
    owner[200:400] = 1
    owner[400:500] = 2
    owner[500:550] = 0xFFFFFFFF        # free particles
    owner[550:600] = 0                 # claim body 0 but lie outside its slice
    pos = np.concatenate([rest, rest + 2.0, rest - 2.0]).astype(np.float32)
    s = O.Solver(600, 4, precision="f32")
    s.set_bodies([dict(first=0, count=200), dict(first=200, count=200, flags=1), dict(first=400, count=100, flags=2)])
    s.set_particles(pos, np.concatenate([rest, rest, rest]), owner)
    before = s.planes.copy()
    for _ in range(5):
        s.step()
    after = s.planes
    assert np.all(after[:, 200:400] == before[:, 200:400]), "kinematic body moved"
    assert np.all(after[:, 500:600] == before[:, 500:600]), "unowned particles touched"
    assert (after[1, :200] < before[1, :200]).all(), "body 0 did not fall"
    assert np.abs(after[1, 400:500] - before[1, 400:500]).max() < 1e-6, "weightless body fell"


def test_full_plasticity_makes_the_deformation_the_rest_shape():
    rest = ball(256, 0.5, 5)
    squash = rest * np.array([1.3, 0.6, 1.0], dtype=np.float32)
    s = O.Solver(256, 1, gravity=(0, 0, 0), substeps=1, precision="f64")
    s.set_bodies([dict(first=0, count=256, stiffness=0.0, plasticity=1.0, flags=2)])
    s.set_particles(squash, rest, np.zeros(256, np.uint32))
    s.step()
    # Stiffness 0 keeps the particles; plasticity 1 rebuilds rest = R^T (p - com) + restCom.
    q = s.state[0, 6:10]
    R = quat_to_mat(q)
    com, rcom = s.state[0, 0:3], s.state[0, 3:6]
    want = (s.get(0)[:256] - com) @ R + rcom
    assert np.allclose(s.get(4)[:256], want, atol=1e-12)
