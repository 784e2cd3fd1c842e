"""The oracle's SurfaceExtractor (lib/src/sim/extract.ts:85-484), pinned on the CPU.

* an independent numpy restatement of naive surface nets (corner distances of an
  exact sphere, edge crossings, averaging, one snap along the central-difference
  normal) gives the same vertices in the same slots, in f64;
* the quads form a closed, consistently wound 2-manifold: every edge is shared by
  two triangles, V - E + F = 2, and every triangle faces away from the centre;
* capacity: a cut-down extraction keeps the first slots of the full one, and
  exactly the quads whose four vertices survived, in the same order;
* the body record (first, count, centroid, identity rotation), the motion snapshot
  and emitted velocity, and the material of analytic and volume leaves.
"""
import numpy as np
import pytest

import oracle as O
from clayhip.brush import make_brush

CENTRE = np.array([0.1, 0.05, -0.02])
RADIUS = 0.7
# Brush records hold f32: the sphere the oracle sees.
C32 = CENTRE.astype(np.float32).astype(np.float64)
R32 = float(np.float32(RADIUS))


def sphere_collider(material: float = 0.0):
    return O.Collider().add_analytic([make_brush({"kind": "sphere", "size": RADIUS, "pos": tuple(CENTRE)})],
                                     material=material)


def run(res=20, capacity=3000, precision="f64", origin=(-1.0, -1.0, -1.0), size=2.0, col=None, motion=False):
    s = O.Solver(4096, 2, precision=precision)
    b = s.add_body()
    e = O.Extractor(s, col or sphere_collider(), resolution=res, body=b, base=0, capacity=capacity, motion=motion)
    e.set_region(origin, size)
    e.extract()
    return s, e


def numpy_surface_nets(res, origin, size):
    """Surface nets of the exact sphere distance, written independently of the oracle."""
    c = size / res
    g = np.arange(res)
    k, j, i = np.meshgrid(g, g, g, indexing="ij")          # flat index = i + res (j + res k)
    p0 = np.asarray(origin)[None] + np.stack([i.ravel(), j.ravel(), k.ravel()], -1) * c
    dist = lambda p: np.linalg.norm(p - C32, axis=-1) - R32
    corners = {}
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                q = p0 + np.array([dx, dy, dz]) * c
                corners[(dx, dy, dz)] = (q, dist(q))
    edges = [((0, 0, 0), (1, 0, 0)), ((0, 1, 0), (1, 1, 0)), ((0, 0, 1), (1, 0, 1)), ((0, 1, 1), (1, 1, 1)),
             ((0, 0, 0), (0, 1, 0)), ((1, 0, 0), (1, 1, 0)), ((0, 0, 1), (0, 1, 1)), ((1, 0, 1), (1, 1, 1)),
             ((0, 0, 0), (0, 0, 1)), ((1, 0, 0), (1, 0, 1)), ((0, 1, 0), (0, 1, 1)), ((1, 1, 0), (1, 1, 1))]
    acc = np.zeros((p0.shape[0], 3))
    w = np.zeros(p0.shape[0])
    for a, b in edges:
        (pa, da), (pb, db) = corners[a], corners[b]
        cross = da * db < 0
        with np.errstate(divide="ignore", invalid="ignore"):
            t = np.clip(da / (da - db), 0, 1)
        pt = pa * (1 - t)[:, None] + pb * t[:, None]
        acc += np.where(cross[:, None], pt, 0)
        w += cross
    has = w >= 0.5
    v = acc[has] / w[has, None]
    e = 1e-3
    grad = np.stack([dist(v + e * ax) - dist(v - e * ax) for ax in np.eye(3)], -1)
    n = grad / np.linalg.norm(grad, axis=1, keepdims=True)
    p = v - n * dist(v)[:, None]
    return np.flatnonzero(has), p, n


def test_vertices_match_an_independent_restatement():
    res, origin, size = 20, (-1.0, -1.0, -1.0), 2.0
    s, e = run(res, precision="f64", origin=origin, size=size)
    cells, p, n = numpy_surface_nets(res, origin, size)
    nv = int(e.counts[0])
    assert nv == cells.size
    assert np.array_equal(np.flatnonzero(e.ids), cells)
    assert np.array_equal(e.ids[cells], np.arange(1, nv + 1))   # slots in ascending cell index
    assert np.allclose(s.get(0)[:nv], p, atol=1e-12)
    assert np.allclose(s.aux[:3, :nv].T, n, atol=1e-12)
    # The exact distance makes the snap land on the sphere.
    assert np.abs(np.linalg.norm(p - C32, axis=1) - R32).max() < 1e-9


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_quads_close_a_consistently_wound_surface(precision):
    s, e = run(24, precision=precision)
    nv, nq = int(e.counts[0]), int(e.counts[1])
    assert e.draw[0] == 6 * nq and e.draw[1] == 1
    tri = e.indices[:6 * nq].reshape(-1, 3).astype(np.int64)
    assert tri.max() < nv
    edges = np.sort(np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]]), axis=1)
    _, mult = np.unique(edges, axis=0, return_counts=True)
    assert (mult == 2).all(), "an edge is not shared by exactly two triangles"
    # Directed edges: each appears once, so neighbouring triangles agree on winding.
    directed = np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]])
    assert np.unique(directed, axis=0).shape[0] == directed.shape[0]
    # Quads: V - E + F = 2 with E = 2F for a closed quad mesh.
    assert nv - 2 * nq + nq == 2
    p = s.get(0)[:nv].astype(np.float64)
    a, b, c = p[tri[:, 0]], p[tri[:, 1]], p[tri[:, 2]]
    normal = np.cross(b - a, c - a)
    out = (a + b + c) / 3 - CENTRE
    assert ((normal * out).sum(1) > 0).all(), "a triangle faces inwards"


def test_body_record_after_extraction():
    s, e = run(16, precision="f64")
    nv = int(e.counts[0])
    assert tuple(s.body3[0]) == (0, nv, 0)
    rest = s.get(4)[:nv]
    assert np.allclose(s.state[0, 0:3], rest.mean(0), atol=1e-12)
    assert np.allclose(s.state[0, 3:6], rest.mean(0), atol=1e-12)
    assert tuple(s.state[0, 6:10]) == (0, 0, 0, 1)
    assert (s.owner[:nv] == 0).all() and (s.owner[nv:] == 0xFFFFFFFF).all()
    assert np.array_equal(s.get(1)[:nv], s.get(0)[:nv]) and np.array_equal(rest, s.get(0)[:nv])


def test_capacity_keeps_the_first_slots_and_their_quads():
    full_s, full = run(20, precision="f32")
    cut_s, cut = run(20, capacity=200, precision="f32")
    assert cut.counts[0] == full.counts[0]                     # found, before the cut
    assert np.array_equal(cut_s.get(0)[:200], full_s.get(0)[:200])
    assert cut_s.body3[0, 1] == 200
    ftri = full.indices[:6 * int(full.counts[1])].reshape(-1, 6)
    keep = ftri[(ftri < 200).all(1)]
    assert int(cut.counts[1]) == keep.shape[0]
    kept = min(int(cut.counts[1]), cut.k.quad_capacity)
    assert np.array_equal(cut.indices[:6 * kept].reshape(-1, 6), keep[:kept])
    assert cut.draw[0] == 6 * kept


def test_motion_snapshot_moves_the_lattice_and_sets_the_velocity():
    s, e = run(16, precision="f64", motion=True, origin=(-1.0, -1.0, -1.0))
    nv = int(e.counts[0])
    first = s.get(0)[:nv].copy()
    shift = np.array([0.3, -0.2, 0.1])
    s.planes[0:3, :nv] += shift[:, None]
    s.planes[6:9, :nv] = np.array([1.0, 2.0, -0.5])[:, None]
    e.set_region((-1.0, -1.0, -1.0), 2.0, preserve_motion=True)
    e.extract()
    centre = first.mean(0) + shift
    assert np.allclose(e.motion[0:3], centre, atol=1e-12)
    assert np.allclose(e.motion[4:7], [1.0, 2.0, -0.5], atol=1e-12)
    # The lattice now starts at origin + centre; the sphere did not move, so the
    # new vertices are the surface nets of a shifted lattice, all on the sphere.
    nv2 = int(e.counts[0])
    p2 = s.get(0)[:nv2]
    assert np.abs(np.linalg.norm(p2 - C32, axis=1) - R32).max() < 1e-9
    assert np.allclose(s.get(2)[:nv2], [1.0, 2.0, -0.5]) and np.allclose(s.get(3)[:nv2], [1.0, 2.0, -0.5])


def test_material_of_analytic_and_volume_leaves():
    s, e = run(12, precision="f32", col=sphere_collider(material=2.0))
    nv = int(e.counts[0])
    assert (s.aux[3, :nv] == 2.0).all()
    # A volume leaf whose material channel reads 5 everywhere, unioned with a far sphere.
    res = 32
    dist = np.zeros((res, res, res), dtype=np.float32)
    g = (np.arange(res) + 0.5) / res * 4.0 - 2.0
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    dist = (np.sqrt(x * x + y * y + z * z) - 0.8) / (4 * 4.0 / res)
    d16 = np.clip(dist, -1, 1).astype(np.float16).view(np.uint16)
    m16 = np.full(res ** 3, 5.0, dtype=np.float16).view(np.uint16)
    far = make_brush({"kind": "sphere", "size": 0.3, "pos": (30.0, 0.0, 0.0)})
    col = O.Collider().add_analytic([far], material=2.0).add_volume(d16, res, (-2, -2, -2), 4.0, 4, mat16=m16)
    s, e = run(16, precision="f32", col=col, origin=(-1.2, -1.2, -1.2), size=2.4)
    nv = int(e.counts[0])
    assert nv > 100 and (s.aux[3, :nv] == 5.0).all()
