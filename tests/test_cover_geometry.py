"""The surface bake's interior radius, checked against brute force on the CPU.

The fluid's density pass classes each neighbour offset d (0.1875 h < |d| < 1.5 h)
by the signs of its components of magnitude at least 0.35 |d| into one of the 26
lattice directions; when every class holds a neighbour closer than dmax (h, 1.25 h or
1.5 h) the particle's squared interior radius is 1.25 dmax^2, and the splat skips
its voxels beyond that radius, because a neighbour is then strictly nearer every
such voxel centre (clay-engine_amd/csrc/fluid.hip, Cover).  The classes and the
radius are restated here; what pins them is the brute-force search for a voxel
centre beyond the radius that no counted neighbour is nearer to, over jittered,
stretched and random neighbourhoods and over directions on the class borders.
The widest pair of directions inside one class needs all but ~9% of the
stated radius.
"""
import numpy as np
import pytest

H = 0.1
DMAX = (1.0, 1.25, 1.5)       # levels, in h
T = 0.35                      # class threshold on |component| / |d|


def classify(d: np.ndarray) -> np.ndarray:
    """Per axis floor(d_a / (2 T |d|) + 1.5): 0, 1 or 2 as d_a / |d| is below -T,
    within [-T, T) or at least T; the bit is s_x + 3 s_y + 9 s_z (Cover::add)."""
    r = np.sqrt((d * d).sum(1))
    s = np.floor(d / (2 * T * r[:, None]) + 1.5).astype(np.int64)
    assert ((s >= 0) & (s <= 2)).all()
    return s[:, 0] + 3 * s[:, 1] + 9 * s[:, 2]


def interior(p: np.ndarray, nbrs: np.ndarray):
    """(squared radius, neighbours that fill the level) or (inf, None)."""
    d = p - nbrs
    r2 = (d * d).sum(1)
    ok = (r2 > 0.03515625 * H * H) & (r2 < 2.25 * H * H)
    for lv in DMAX:
        sel = ok & (r2 < (lv * H) ** 2)
        cls = classify(d[sel])
        if len(set(cls.tolist()) - {13}) == 26:
            return 1.25 * (lv * H) ** 2, nbrs[sel]
    return np.inf, None


def directions(rng, n: int) -> np.ndarray:
    """Random unit vectors plus ones whose components sit on the class thresholds."""
    w = rng.normal(size=(n, 3))
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    edge = rng.normal(size=(n, 3))
    edge /= np.linalg.norm(edge, axis=1, keepdims=True)
    k = rng.integers(0, 3, size=n)
    s = np.sign(edge[np.arange(n), k])
    edge[np.arange(n), k] = s * T * (1 + rng.uniform(-1e-6, 1e-6, size=n))
    rest = [(k + 1) % 3, (k + 2) % 3]
    scale = np.sqrt((1 - (T * T)) / (edge[np.arange(n), rest[0]] ** 2 + edge[np.arange(n), rest[1]] ** 2))
    for r in rest:
        edge[np.arange(n), r] *= scale
    return np.concatenate([w, edge])


def clouds(rng):
    g = np.arange(-4, 5) * 0.05
    zz, yy, xx = np.meshgrid(g, g, g, indexing="ij")
    lattice = np.stack([xx.ravel(), yy.ravel(), zz.ravel()], axis=-1)
    yield "jittered lattice", lattice + rng.uniform(-0.005, 0.005, size=lattice.shape)
    yield "stretched lattice", lattice * 1.3 + rng.uniform(-0.01, 0.01, size=lattice.shape)
    yield "random", rng.uniform(-0.2, 0.2, size=(520, 3))
    yield "random dense", rng.uniform(-0.15, 0.15, size=(900, 3))


def violations(factor: float, rng, max_particles: int = 60) -> tuple[int, int]:
    """Counts (pairs checked, pairs where no counted neighbour is strictly nearer)
    with the interior radius scaled by `factor`."""
    checked = bad = 0
    for _, cloud in clouds(rng):
        centre = np.argsort(np.linalg.norm(cloud, axis=1))[:max_particles]
        for i in centre:
            p = cloud[i]
            r2, nb = interior(p, np.delete(cloud, i, axis=0))
            if not np.isfinite(r2):
                continue
            w = directions(rng, 400)
            r = np.sqrt(r2) * factor * rng.uniform(1.0 + 1e-9, 3.0, size=(len(w), 1))
            c = p + w * r
            dp = ((c - p) ** 2).sum(1)
            dq = ((c[:, None, :] - nb[None, :, :]) ** 2).sum(2).min(1)
            checked += len(c)
            bad += int((dq >= dp * (1 - 1e-5)).sum())
    return checked, bad


def test_every_direction_has_a_class_within_30_5_degrees():
    rng = np.random.default_rng(0)
    w = directions(rng, 500_000)
    cls = classify(w)
    assert not np.any(cls == 13)
    s = np.stack([(cls % 3) - 1, ((cls // 3) % 3) - 1, (cls // 9) - 1], axis=-1).astype(np.float64)
    a = s / np.linalg.norm(s, axis=1, keepdims=True)
    ang = np.degrees(np.arccos(np.clip((w * a).sum(1), -1, 1)))
    assert ang.max() < 30.5


def test_interior_radius_is_dominated_beyond_it():
    checked, bad = violations(1.0, np.random.default_rng(1))
    assert checked > 20_000
    assert bad == 0


def widest_pairs(rng):
    """Per class, two sampled directions of the class at the widest angle."""
    v = rng.normal(size=(400_000, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    cls = classify(v)
    pairs = []
    for k in range(27):
        if k == 13:
            continue
        m = v[cls == k]
        axis = m.mean(0) / np.linalg.norm(m.mean(0))
        u = m[np.argmin(m @ axis)]
        pairs.append((u, m[np.argmin(m @ u)]))
    return pairs


def test_the_stated_radius_is_needed_for_the_widest_pair_of_a_class():
    """One neighbour q at 0.999 dmax in one direction of a class, a voxel centre in
    the direction of the class farthest from it: q is nearer at the stated radius
    (1.118 dmax) for every class, and no longer for some class at 0.91 of it, so
    the class angle and the factor are each within a few percent of what they
    must be."""
    pairs = widest_pairs(np.random.default_rng(3))
    assert max(np.degrees(np.arccos(u @ w)) for u, w in pairs) > 60.5
    dmax = H
    radius = np.sqrt(1.25) * dmax

    def dominated(f):
        out = []
        for u, w in pairs:
            q = u * 0.999 * dmax
            c = w * radius * f
            out.append(((c - q) ** 2).sum() < (c @ c) * (1 - 1e-5))
        return np.array(out)

    assert dominated(1.0).all()
    assert not dominated(0.91).all()


@pytest.mark.parametrize("axis", range(3))
def test_a_particle_on_a_free_surface_is_not_interior(axis):
    g = np.arange(-4, 5) * 0.05
    zz, yy, xx = np.meshgrid(g, g, g, indexing="ij")
    cloud = np.stack([xx.ravel(), yy.ravel(), zz.ravel()], axis=-1)
    cloud = cloud[cloud[:, axis] <= 0]  # particles on the plane x_axis = 0 have nothing above
    p = np.zeros(3)
    others = cloud[np.linalg.norm(cloud, axis=1) > 0]
    assert np.isinf(interior(p, others)[0])
    q = np.zeros(3)
    q[axis] = -0.1  # two layers down: covered within h
    others = cloud[np.linalg.norm(cloud - q, axis=1) > 0]
    assert np.isfinite(interior(q, others)[0])
