"""The oracle's tracer reads of a field (lib/src/trace/field.ts:20-315), pinned on the CPU.

* sample / field of a volume at every mip (and above its top mip) against an
  independent numpy trilinear, clamp-to-edge, of that level's texels, times
  band * voxel * 2^m; outside the box, (distance to the box, 1) or (.., 0);
* analytic leaves: (v, clamp(v / band, -1, 1)) and (v, material) for a sphere's
  exact distance;
* unions return the whole pair of the nearer child;
* entry against the textbook slab intersection of a ray with the box grown by the
  radius (aperture 0), NEVER for misses, tMin for rays starting inside that are
  still inside at tMin, and an aperture that can only bring the entry earlier.
"""
import numpy as np
import pytest

import oracle as O
from clayhip.brush import make_brush

RES, MIPS, WORLD, BAND = 32, 3, 4.0, 4.0
ORIGIN = np.array([-2.0, -1.0, -2.0])


def levels(seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for m in range(MIPS):
        r = RES >> m
        d = rng.uniform(-1, 1, size=(r, r, r)).astype(np.float16).view(np.uint16)
        mt = rng.uniform(0, 8, size=(r, r, r)).astype(np.float16).view(np.uint16)
        out.append((d, mt))
    return out


def volume_collider(lv):
    return O.Collider().add_volume(lv[0][0], RES, tuple(ORIGIN), WORLD, BAND, mat16=lv[0][1], levels=lv)


def trilinear(tex16, p, m):
    r = RES >> m
    t = tex16.view(np.float16).astype(np.float64)          # [z, y, x]
    u = (p - ORIGIN) / WORLD * r - 0.5
    f0 = np.floor(u)
    fr = u - f0
    i0 = f0.astype(np.int64)
    acc = np.zeros(p.shape[0])
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                w = (np.where(dx, fr[:, 0], 1 - fr[:, 0]) * np.where(dy, fr[:, 1], 1 - fr[:, 1])
                     * np.where(dz, fr[:, 2], 1 - fr[:, 2]))
                ix = np.clip(i0[:, 0] + dx, 0, r - 1)
                iy = np.clip(i0[:, 1] + dy, 0, r - 1)
                iz = np.clip(i0[:, 2] + dz, 0, r - 1)
                acc += w * t[iz, iy, ix]
    return acc


def sd_box(p, c, h):
    q = np.abs(p - c) - h
    return np.linalg.norm(np.maximum(q, 0), axis=1) + np.minimum(q.max(1), 0)


@pytest.mark.parametrize("mip", [0, 1, 2, 3, 7])
def test_volume_sample_and_field_at_every_level(mip):
    lv = volume_collider(levels())
    rng = np.random.default_rng(mip)
    p = ORIGIN + rng.uniform(-0.3, WORLD + 0.3, size=(3000, 3))
    s = lv.query(p, mip, "sample", "f64")
    f = lv.query(p, mip, "field", "f64")
    m = min(mip, MIPS - 1)
    voxel = WORLD / RES
    bw = BAND * voxel * 2 ** m
    dbox = sd_box(p, ORIGIN + WORLD / 2, WORLD / 2)
    inside = dbox <= 0
    assert inside.sum() > 1000 and (~inside).sum() > 100
    raw = trilinear(levels()[m][0], p[inside], m)
    mat = trilinear(levels()[m][1], p[inside], m)
    assert np.allclose(s[inside, 1], raw, atol=1e-12) and np.allclose(s[inside, 0], raw * bw, atol=1e-12)
    assert np.allclose(f[inside, 0], raw * bw, atol=1e-12) and np.allclose(f[inside, 1], mat, atol=1e-12)
    assert np.allclose(s[~inside, 0], dbox[~inside], atol=1e-12) and (s[~inside, 1] == 1).all()
    assert np.allclose(f[~inside, 0], dbox[~inside], atol=1e-12) and (f[~inside, 1] == 0).all()


def test_analytic_and_union_pairs():
    sph = make_brush({"kind": "sphere", "size": 0.6, "pos": (0.3, 0.2, -0.1)})
    a = O.Collider().add_analytic([sph], material=5.0, band=2.0)
    rng = np.random.default_rng(1)
    p = rng.uniform(-3, 3, size=(2000, 3))
    c32 = np.array([0.3, 0.2, -0.1], np.float32).astype(np.float64)
    v = np.linalg.norm(p - c32, axis=1) - float(np.float32(0.6))
    s = a.query(p, 2, "sample", "f64")
    f = a.query(p, 0, "field", "f64")
    assert np.allclose(s[:, 0], v, atol=1e-12) and np.allclose(s[:, 1], np.clip(v / 2.0, -1, 1), atol=1e-12)
    assert np.allclose(f[:, 0], v, atol=1e-12) and (f[:, 1] == 5.0).all()
    lv = levels(3)
    u = O.Collider().add_analytic([sph], material=5.0, band=2.0).add_volume(
        lv[0][0], RES, tuple(ORIGIN), WORLD, BAND, mat16=lv[0][1], levels=lv)
    vo = volume_collider(lv)
    for mip in (0, 1):
        su, sa, sv = (c.query(p, mip, "sample", "f64") for c in (u, a, vo))
        pick_a = sa[:, 0] < sv[:, 0]
        want = np.where(pick_a[:, None], sa, sv)
        assert np.array_equal(su, want)
        assert pick_a.any() and (~pick_a).any()


def slab(ro, rd, lo, hi):
    """Textbook ray/box slab test in f64: (t_near, t_far)."""
    with np.errstate(divide="ignore", invalid="ignore"):
        t1 = (lo - ro) / rd
        t2 = (hi - ro) / rd
    tn = np.nanmax(np.minimum(t1, t2), axis=1)
    tf = np.nanmin(np.maximum(t1, t2), axis=1)
    return tn, tf


def test_entry_against_the_slab_intersection():
    lv = levels()
    c = volume_collider(lv)
    rng = np.random.default_rng(5)
    n = 3000
    ro = rng.uniform(-8, 8, size=(n, 3))
    rd = rng.normal(size=(n, 3))
    rd /= np.linalg.norm(rd, axis=1, keepdims=True)
    tmin, radius = 0.05, 0.3
    rays = np.concatenate([ro, rd, np.full((n, 1), tmin), np.full((n, 1), radius), np.zeros((n, 1))], axis=1)
    t = c.entry(rays, "f64")
    lo, hi = ORIGIN - radius, ORIGIN + WORLD + radius
    tn, tf = slab(ro, rd, lo, hi)
    want = np.where(tf >= np.maximum(tn, tmin), np.maximum(tn, tmin), 1e9)
    assert np.allclose(t, want, rtol=1e-9, atol=1e-9)
    assert (t == 1e9).sum() > 100 and (t < 1e9).sum() > 100
    # Starting inside, the entry is tMin unless the ray leaves the box before tMin.
    inside = np.all((ro > lo) & (ro < hi), axis=1) & (tf >= tmin)
    assert inside.any() and np.allclose(t[inside], tmin)
    # A positive aperture grows the box along the ray: the entry can only come earlier.
    rays_a = rays.copy()
    rays_a[:, 8] = 0.2
    ta = c.entry(rays_a, "f64")
    assert (ta <= t + 1e-12).all() and (ta < t).any()


def test_offset_reads_the_child_at_the_moved_point():
    """offsetField (field.ts:388-429): every read of the offset field at p is the
    child's read at the f32 point p - o, bit for bit."""
    lv = levels(7)
    o = np.array([0.37, -0.21, 1.13], np.float32)
    moved = volume_collider(lv).offset(o)
    plain = volume_collider(lv)
    rng = np.random.default_rng(8)
    p = (ORIGIN + rng.uniform(-0.5, WORLD + 1.5, size=(2000, 3))).astype(np.float32)
    q = (p - o).astype(np.float32)
    for mip in (0, 1, 2):
        for what in ("sample", "field"):
            assert np.array_equal(moved.query(p, mip, what, "f32"), plain.query(q, mip, what, "f32"))
    d0, n0 = moved.eval(p, "f32")
    d1, n1 = plain.eval(q, "f32")
    assert np.array_equal(d0, d1) and np.array_equal(n0, n1)
    rays = np.concatenate([rng.uniform(-8, 8, size=(500, 3)), rng.normal(size=(500, 3)),
                           np.full((500, 1), 0.05), np.full((500, 1), 0.2), np.zeros((500, 1))], axis=1).astype(np.float32)
    shifted = rays.copy()
    shifted[:, :3] = (rays[:, :3] - o).astype(np.float32)
    assert np.array_equal(moved.entry(rays, "f32"), plain.entry(shifted, "f32"))


def test_lerp_of_a_morph_pair():
    """lerpField (field.ts:316-386): sample mixes the distance and keeps the smaller band
    value, field mixes both, the normal is the normalized mix, entry the nearer one;
    t is clamped, so t <= 0 and t >= 1 give the two volumes back exactly."""
    la, lb = levels(11), levels(12)
    rng = np.random.default_rng(9)
    p = (ORIGIN + rng.uniform(-0.3, WORLD + 0.3, size=(2000, 3))).astype(np.float32)
    for t in (0.3, 0.75):
        lerp = O.Collider().add_lerp(volume_collider(la), volume_collider(lb), t)
        k = np.float32(t)
        for mip in (0, 1, 2):
            a = volume_collider(la).query(p, mip, "sample", "f32")
            b = volume_collider(lb).query(p, mip, "sample", "f32")
            s = lerp.query(p, mip, "sample", "f32")
            assert np.array_equal(s[:, 0], a[:, 0] * (np.float32(1) - k) + b[:, 0] * k)
            assert np.array_equal(s[:, 1], np.minimum(a[:, 1], b[:, 1]))
            fa = volume_collider(la).query(p, mip, "field", "f32")
            fb = volume_collider(lb).query(p, mip, "field", "f32")
            assert np.array_equal(lerp.query(p, mip, "field", "f32"), fa * (np.float32(1) - k) + fb * k)
        _, na = volume_collider(la).eval(p, "f32")
        _, nb = volume_collider(lb).eval(p, "f32")
        _, nl = lerp.eval(p, "f32")
        g = na * (np.float32(1) - k) + nb * k
        assert np.allclose(nl, g / np.linalg.norm(g, axis=1, keepdims=True), atol=2e-6)
    for t, src in ((-0.5, la), (0.0, la), (1.0, lb), (1.7, lb)):
        lerp = O.Collider().add_lerp(volume_collider(la), volume_collider(lb), t)
        assert np.array_equal(lerp.query(p, 1, "sample", "f32")[:, 0], volume_collider(src).query(p, 1, "sample", "f32")[:, 0])
    rays = np.concatenate([rng.uniform(-8, 8, size=(500, 3)), rng.normal(size=(500, 3)),
                           np.full((500, 1), 0.05), np.full((500, 1), 0.2), np.zeros((500, 1))], axis=1).astype(np.float32)
    lerp = O.Collider().add_lerp(volume_collider(la), volume_collider(lb), 0.5)
    assert np.array_equal(lerp.entry(rays, "f32"), volume_collider(la).entry(rays, "f32"))
