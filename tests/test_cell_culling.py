"""The neighbour-cell culling of the density and forces passes is conservative.

Per row (dy, dz) of the 27-cell neighbourhood, a lane drops neighbour cell dx when
the gap between its particle and the cell's box, squared, is at least reach^2 times a
margin (1.0001 in the FP16-LDS kernels, 1.002 in the exact ones).  The exact kernels
rely on this to stay bit-identical to the oracle: a dropped cell must never hold a
particle whose computed squared distance is below reach^2.  This restates the rule
(row_keep in fluid.hip) in f32 numpy with the kernels' cell assignment
(floor((p - origin) / h) through a multiply by 1/h) and checks, by brute force over
random particles placed at cell faces and corners, that every pair with
r^2 < reach^2 (computed in f32) falls in a kept cell.
"""
import numpy as np
import pytest

F = np.float32


def cell_of(p, go, inv_h):
    return np.floor((p - go) * inv_h).astype(np.int64)


def cell_rel(p, b, go, h):
    return (p - (go + (b.astype(F) + F(0.5)) * h)).astype(F)


def row_keep(oi, dy, dz, h, reach2, margin):
    hh = F(0.5) * h
    gy = np.where(dy == 0, F(0), np.maximum(np.where(dy < 0, oi[:, 1] + hh, hh - oi[:, 1]), F(0))).astype(F)
    gz = np.where(dz == 0, F(0), np.maximum(np.where(dz < 0, oi[:, 2] + hh, hh - oi[:, 2]), F(0))).astype(F)
    gyz = (gy * gy + gz * gz).astype(F)
    gl = np.maximum(oi[:, 0] + hh, F(0)).astype(F)
    gr = np.maximum(hh - oi[:, 0], F(0)).astype(F)
    lim = F(reach2 * F(margin))
    return np.stack([gl * gl + gyz < lim, gyz < lim, gr * gr + gyz < lim], axis=-1)


def points_near_faces(rng, n, go, h, cells):
    """Particles on or within a few ulps of cell faces, edges and corners, and uniform ones."""
    base = rng.integers(1, cells - 1, size=(n, 3)).astype(F)
    frac = rng.uniform(0, 1, size=(n, 3)).astype(F)
    snap = rng.random((n, 3)) < 0.5
    frac = np.where(snap, rng.choice(np.array([0, 1e-7, 1 - 1e-7, 0.5], F), size=(n, 3)), frac).astype(F)
    p = (go + (base + frac) * h).astype(F)
    jitter = rng.integers(-3, 4, size=(n, 3))
    return np.nextafter(p, p + np.sign(jitter).astype(F) * F(1)) if rng.random() < 0.5 else p


def count_violations(reach_h, margin, seed):
    rng = np.random.default_rng(seed)
    h = F(0.1)
    inv_h = F(1) / h
    go = np.array([-12.1, -4.1, -12.1], F)
    reach2 = F(reach_h * reach_h) * (h * h)
    cells = 6
    pi = points_near_faces(rng, 3000, go, h, cells)
    pj = points_near_faces(rng, 3000, go, h, cells)
    bi, bj = cell_of(pi, go, inv_h), cell_of(pj, go, inv_h)
    oi = cell_rel(pi, bi, go, h)
    violations = 0
    checked = 0
    for a in range(0, pi.shape[0], 500):
        P, B, O = pi[a:a + 500], bi[a:a + 500], oi[a:a + 500]
        d = (P[:, None, :] - pj[None, :, :]).astype(F)
        r2 = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1] + d[..., 2] * d[..., 2]).astype(F)
        rel = bj[None, :, :] - B[:, None, :]
        near = (r2 < reach2) & (np.abs(rel) <= 1).all(-1)
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                keep = row_keep(O, dy, dz, h, reach2, margin)           # (rows, 3)
                in_row = near & (rel[..., 1] == dy) & (rel[..., 2] == dz)
                for dx in (-1, 0, 1):
                    sel = in_row & (rel[..., 0] == dx)
                    checked += int(sel.sum())
                    violations += int((sel & ~keep[:, dx + 1][:, None]).sum())
    return checked, violations


@pytest.mark.parametrize("reach_h,margin", [(1.0, 1.0001), (1.0, 1.002), (1.5, 1.0001), (1.5, 1.002)])
def test_no_neighbour_within_reach_in_a_dropped_cell(reach_h, margin):
    checked, violations = count_violations(reach_h, margin, int(reach_h * 10 + margin * 1e4) % 2**31)
    assert checked > 10000
    assert violations == 0


def test_the_check_catches_a_rule_that_drops_too_much():
    """A margin below 1 (a rule that drops cells holding neighbours) must be caught."""
    _, violations = count_violations(1.0, 0.95, 7)
    assert violations > 0


def test_culling_drops_a_quarter_of_the_cells_at_reach_h():
    """Particles uniform in their cells: about 1 - (1 + 6 + 12 pi/4 + 8 pi/6) / 27 of
    the neighbour cells lie beyond h."""
    rng = np.random.default_rng(3)
    h = F(0.1)
    oi = (rng.uniform(-0.5, 0.5, size=(200000, 3)) * h).astype(F)
    kept = 0
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            kept += row_keep(oi, dy, dz, h, h * h, 1.0001).sum(-1).mean()
    want = 1 + 6 + 12 * np.pi / 4 + 8 * np.pi / 6
    assert abs(kept - want) < 0.05
