"""An independent numpy f64 restatement of the FluidSim passes, written from the
reference shaders, used to check the C oracle.

It shares nothing with the oracle: neighbours come from all pairs (no grid, no
bucket scan), sums run in numpy's own order, and the collider is its own Quilez
round-box plus central differences.  The passes follow
  density    lib/src/sim/fluid.ts:333-368
  forces     lib/src/sim/fluid.ts:370-439
  integrate  lib/src/sim/fluid.ts:444-491
  normal     lib/src/trace/field.ts:225-234 (analyticField.normal)
"""
from __future__ import annotations

import math

import numpy as np


def constants(spacing: float, h: float, density: float = 1.0, substeps: int = 2) -> dict:
    """fluid.ts:164-188: kernel constants, particle mass and the lattice rest density."""
    poly6 = 315.0 / (64.0 * math.pi * h ** 9)
    spiky = -45.0 / (math.pi * h ** 6)
    mass = density * spacing ** 3
    reach = math.ceil(h / spacing)
    rest = 0.0
    for z in range(-reach, reach + 1):
        for y in range(-reach, reach + 1):
            for x in range(-reach, reach + 1):
                r2 = (x * x + y * y + z * z) * spacing * spacing
                if r2 < h * h:
                    rest += mass * poly6 * (h * h - r2) ** 3
    return {"h": h, "poly6": poly6, "spiky": spiky, "mass": mass, "rest": rest, "dt": 1.0 / 60.0 / substeps}


def _pairs(pos: np.ndarray):
    d = pos[:, None, :] - pos[None, :, :]  # d[i, j] = p_i - p_j
    r2 = (d * d).sum(-1)
    return d, r2


def density(pos: np.ndarray, k: dict, stiffness: float):
    """rho_i = sum over j with r^2 < h^2 (self included) of m poly6 (h^2 - r^2)^3;
    stored as max(rho, 0.1 rho0), pressure max(k (rho - rho0), 0) from the raw rho."""
    h2 = k["h"] * k["h"]
    _, r2 = _pairs(pos)
    w = np.where(r2 < h2, (h2 - r2) ** 3, 0.0)
    rho = k["mass"] * k["poly6"] * w.sum(1)
    return np.maximum(rho, k["rest"] * 0.1), np.maximum(stiffness * (rho - k["rest"]), 0.0)


def forces(pos, vel, rho, pres, k: dict, gravity, viscosity: float, damping: float, max_speed: float):
    """Spiky pressure gradient plus the XSPH blend, then gravity, damping and the
    speed cap: the new velocity nvel of every particle."""
    h = k["h"]
    h2 = h * h
    n = pos.shape[0]
    d, r2 = _pairs(pos)
    near = (r2 < h2) & ~np.eye(n, dtype=bool)
    w = np.where(near, (h2 - r2) ** 3, 0.0)
    vsum = vel * h ** 6 + w @ vel
    wsum = h ** 6 + w.sum(1)
    r = np.sqrt(r2)
    live = near & (r > 1e-6)
    rs = np.where(live, r, 1.0)
    coef = np.where(live, k["mass"] * (pres[:, None] + pres[None, :]) * 0.5 / rho[None, :] * k["spiky"]
                    * (h - rs) ** 2 / rs, 0.0)
    fpress = -(coef[:, :, None] * d).sum(1)
    acc = fpress / rho[:, None] + np.asarray(gravity, dtype=np.float64)
    t = min(max(viscosity, 0.0), 1.0)
    v = vel * (1.0 - t) + (vsum / wsum[:, None]) * t
    v = (v + acc * k["dt"]) * damping
    speed = np.linalg.norm(v, axis=1)
    cap = speed > max_speed
    v[cap] *= (max_speed / speed[cap])[:, None]
    return v


def round_box(p: np.ndarray, centre, size, radius: float) -> np.ndarray:
    """Quilez's rounded box: |max(q, 0)| + min(max(q), 0) - r with q = |p - c| - size + r."""
    q = np.abs(p - np.asarray(centre, dtype=np.float64)) - np.asarray(size, dtype=np.float64) + radius
    return np.linalg.norm(np.maximum(q, 0.0), axis=-1) + np.minimum(q.max(-1), 0.0) - radius


def central_normal(sdf, p: np.ndarray, eps: float) -> np.ndarray:
    """analyticField.normal: central differences at +-eps per axis, normalised, or
    straight up when the difference vector is shorter than 1e-9."""
    g = np.empty_like(p)
    for a in range(3):
        e = np.zeros(3)
        e[a] = eps
        g[:, a] = sdf(p + e) - sdf(p - e)
    ln = np.linalg.norm(g, axis=1)
    out = np.where((ln > 1e-9)[:, None], g / np.where(ln > 1e-9, ln, 1.0)[:, None], np.array([0.0, 1.0, 0.0]))
    return out


def integrate(pos, nvel, dt: float, radius: float, contact_depth: float, sdf, normal, wall_lo, wall_hi):
    """Advance by nvel, push out of the collider along its normal and drop the
    inward velocity, note contacts (in ascending id), clamp to the walls and zero the
    velocity of every clamped component.  Returns (pos, vel, contact ids, points)."""
    v = nvel.copy()
    p = pos + v * dt
    dw = sdf(p)
    hit = dw < radius
    nr = normal(p[hit])
    p[hit] = p[hit] + nr * (radius - dw[hit])[:, None]
    vn = np.minimum((v[hit] * nr).sum(1), 0.0)
    v[hit] = v[hit] - nr * vn[:, None]
    ids = np.nonzero(hit)[0]
    rec = ids[dw[hit] < contact_depth]
    points = np.concatenate([p[rec], np.linalg.norm(v[rec], axis=1)[:, None]], axis=1)
    cl = np.minimum(np.maximum(p, np.asarray(wall_lo)), np.asarray(wall_hi))
    v = np.where(cl == p, v, 0.0)
    return cl, v, rec, points
