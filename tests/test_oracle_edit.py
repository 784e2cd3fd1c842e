"""Pins the oracle's SdfEditor / TileGrid / MipRefiner / body-centroid functions.

Each oracle function is checked against an independent numpy formulation
written from the reference's description (lib/src/field/modify.ts:98-238,
tilegrid.ts:207-242, mips.ts:78-177, sim/pbd.ts:240-266), against closed forms
where the mathematics fixes the answer, and, for the erosion edits, against
the brush record BrushSet.make builds (brush.ts:539-564).
"""
import numpy as np
import pytest

from clayhip import make_brush

f32 = np.float32


def sphere_edit(pos, r, op="cut", smooth=0.02, material=0.0):
    return make_brush({"kind": "sphere", "size": r, "pos": pos, "op": op, "smooth": smooth, "material": material})


# ---- TileGrid bin (modify.ts:98-128) --------------------------------------------
def numpy_bin(res, world, origin, band, capacity, edits):
    tr = res // 8
    voxel = world / res
    tsize, thalf, bw = f32(voxel * 8), f32(voxel * 8 / 2), f32(band * voxel)
    org = np.array(origin, dtype=np.float64).astype(f32)
    t = np.arange(tr, dtype=f32)
    tz, ty, tx = np.meshgrid(t, t, t, indexing="ij")
    tc = [org[0] + (tx * tsize + thalf), org[1] + (ty * tsize + thalf), org[2] + (tz * tsize + thalf)]
    counts = np.zeros(tr ** 3, dtype=np.uint32)
    ids = [[] for _ in range(tr ** 3)]
    for b, e in enumerate(edits):
        near = [np.maximum(np.abs(f32(e.pos[a]) - tc[a]) - thalf, f32(0)) for a in range(3)]
        length = np.sqrt((near[0] * near[0] + near[1] * near[1]) + near[2] * near[2])
        hit = (length < f32(e.bound) + bw).reshape(-1)
        for i in np.nonzero(hit)[0]:
            ids[i].append(b)
    for i in range(tr ** 3):
        counts[i] = min(len(ids[i]), capacity)
    return counts, ids


def test_edit_bin_matches_numpy(oracle):
    rng = np.random.default_rng(1)
    res, world, origin, band = 64, 8.0, (-4.0, -2.0, -4.0), 4.0
    edits = [sphere_edit(rng.uniform(-4, 4, 3), float(rng.uniform(0.05, 0.8))) for _ in range(40)]
    edits += [sphere_edit(np.array([0.5, 0.0, 0.5]) + rng.uniform(-0.1, 0.1, 3), 0.3) for _ in range(12)]
    counts, ids = oracle.edit_bin(res, world, origin, band, 8, edits)
    want_counts, want_ids = numpy_bin(res, world, origin, band, 8, edits)
    assert np.array_equal(counts, want_counts)
    for i in np.nonzero(counts)[0]:
        assert list(ids[i][:counts[i]]) == want_ids[i][:8]
    assert counts.max() == 8  # some tile saw more edits than the capacity: truncated in order


# ---- TileGrid dilate (tilegrid.ts:207-242) ---------------------------------------
def test_dilate_matches_window_max(oracle):
    rng = np.random.default_rng(2)
    res, mips = 128, 4
    _, tile_res, offs, total = oracle.volume_derive(res, mips)
    counts = np.zeros(total, dtype=np.uint32)
    fine = (rng.random(tile_res[0] ** 3) < 0.02).astype(np.uint32) * rng.integers(1, 8, tile_res[0] ** 3).astype(np.uint32)
    counts[offs[0]:offs[0] + fine.size] = fine
    for m in range(1, mips):
        oracle.tile_dilate(res, mips, m, counts)
        prev = counts[offs[m - 1]:offs[m - 1] + tile_res[m - 1] ** 3].reshape((tile_res[m - 1],) * 3) != 0
        # independent form: pad by one tile, then the max over the 4^3 window [2c-1, 2c+2]
        pad = np.pad(prev, ((1, 1),) * 3)
        tr = tile_res[m]
        want = np.zeros((tr,) * 3, dtype=bool)
        for dz in range(4):
            for dy in range(4):
                for dx in range(4):
                    want |= pad[dz:dz + 2 * tr:2, dy:dy + 2 * tr:2, dx:dx + 2 * tr:2]
        got = counts[offs[m]:offs[m] + tr ** 3].reshape((tr,) * 3) != 0
        assert np.array_equal(got, want), m


def test_dilate_of_mip0_is_rejected_silently(oracle):
    counts = np.ones(oracle.total_cells(64, 3), dtype=np.uint32)
    before = counts.copy()
    oracle.tile_dilate(64, 3, 0, counts)
    assert np.array_equal(counts, before)


# ---- MipRefiner (mips.ts:78-177) --------------------------------------------------
def child_avg(plane, cx, cy, cz):
    h = f32(0.5)

    def mix(a, b):
        return a * (f32(1) - h) + b * h

    def t(dx, dy, dz):
        return plane[2 * cz + dz, 2 * cy + dy, 2 * cx + dx].astype(np.float16).astype(f32)

    c00 = mix(t(0, 0, 0), t(1, 0, 0))
    c10 = mix(t(0, 1, 0), t(1, 1, 0))
    c01 = mix(t(0, 0, 1), t(1, 0, 1))
    c11 = mix(t(0, 1, 1), t(1, 1, 1))
    return mix(mix(c00, c10), mix(c01, c11))


def numpy_refresh(fine_d, fine_m, band, safety, live_tiles):
    """Vectorised restatement: fine planes are float16 arrays [z, y, x]."""
    fr = fine_d.shape[0]
    mr = fr // 2
    step = f32((1.0 / band) * safety)
    out_d = np.full((mr,) * 3, np.nan, dtype=np.float16)
    out_m = np.full((mr,) * 3, np.nan, dtype=np.float16)
    E, B = 12, 2
    for (tz, ty, tx) in live_tiles:
        o = np.array([tz, ty, tx]) * 8
        idx = [np.clip(o[a] + np.arange(E) - B, 0, mr - 1) for a in range(3)]
        gz, gy, gx = np.meshgrid(*idx, indexing="ij")
        g = child_avg(fine_d, gx, gy, gz) * f32(0.5)
        for k in range(B):
            src = g
            dst = src.copy()
            mg = k + 1
            core = (slice(mg, E - mg),) * 3
            lo = np.full(src[core].shape, f32(-1e9), dtype=f32)
            hi = np.full(src[core].shape, f32(1e9), dtype=f32)
            for dz in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for dx in (-1, 0, 1):
                        if dx == dy == dz == 0:
                            continue
                        w = np.sqrt(f32(dx * dx + dy * dy + dz * dz)) * step
                        nb = src[mg + dz:E - mg + dz, mg + dy:E - mg + dy, mg + dx:E - mg + dx]
                        hi = np.minimum(hi, nb + w)
                        lo = np.maximum(lo, nb - w)
            dst[core] = np.minimum(np.maximum(src[core], lo), hi)
            g = dst
        cz, cy, cx = np.meshgrid(*[o[a] + np.arange(8) for a in range(3)], indexing="ij")
        out_d[o[0]:o[0] + 8, o[1]:o[1] + 8, o[2]:o[2] + 8] = np.clip(g[B:B + 8, B:B + 8, B:B + 8], -1, 1).astype(np.float16)
        out_m[o[0]:o[0] + 8, o[1]:o[1] + 8, o[2]:o[2] + 8] = child_avg(fine_m, cx, cy, cz).astype(np.float16)
    return out_d, out_m


def baked_planes(oracle, res=64, mips=3, world=8.0, origin=(-4.0, -2.0, -4.0), band=4.0):
    brushes = [make_brush({"kind": "roundBox", "size": (3, 0.5, 3), "radius": 0.25, "pos": (0, -1.0, 0),
                           "material": 0.3}),
               make_brush({"kind": "sphere", "size": 1.1, "pos": (0.7, 0.2, -0.4), "material": 0.8})]
    tex = oracle.bake_brushes(res, mips, world, origin, band, brushes)
    return oracle.Planes(res, mips, tex)


def test_mip_refresh_matches_numpy_restatement(oracle):
    planes = baked_planes(oracle)
    res, mips, band, safety = 64, 3, 4.0, 0.98
    _, tile_res, offs, total = oracle.volume_derive(res, mips)
    counts = np.zeros(total, dtype=np.uint32)
    rng = np.random.default_rng(3)
    live = rng.random(tile_res[1] ** 3) < 0.5
    counts[offs[1]:offs[1] + live.size] = live
    fine_d = planes.level(0, "dist").view(np.float16).copy()
    fine_m = planes.level(0, "mat").view(np.float16).copy()
    before = planes.level(1).copy()
    oracle.mip_refresh(planes, 1, band, safety, counts)
    tiles = [(i // (tile_res[1] ** 2), (i // tile_res[1]) % tile_res[1], i % tile_res[1]) for i in np.nonzero(live)[0]]
    want_d, want_m = numpy_refresh(fine_d.astype(f32), fine_m.astype(f32), band, safety, tiles)
    got_d = planes.level(1, "dist").view(np.float16)
    got_m = planes.level(1, "mat").view(np.float16)
    touched = ~np.isnan(want_d.astype(f32))
    assert np.array_equal(got_d[touched].view(np.uint16), want_d[touched].view(np.uint16))
    assert np.array_equal(got_m[touched].view(np.uint16), want_m[touched].view(np.uint16))
    assert np.array_equal(planes.level(1)[~touched], before[~touched])  # dead tiles untouched


def test_mip_refresh_closed_forms(oracle):
    res, mips, band = 32, 2, 4.0
    _, tile_res, offs, total = oracle.volume_derive(res, mips)
    counts = np.zeros(total, dtype=np.uint32)
    counts[offs[1]:] = 1
    # Uniform field: the downsample halves it and the relaxation leaves it alone.
    planes = oracle.Planes(res, mips)
    planes.dist[:] = np.float16(0.5).view(np.uint16)
    planes.mat[:] = np.float16(0.25).view(np.uint16)
    oracle.mip_refresh(planes, 1, band, 0.98, counts)
    assert np.all(planes.level(1, "dist") == np.float16(0.25).view(np.uint16))
    assert np.all(planes.level(1, "mat") == np.float16(0.25).view(np.uint16))
    # A field along x whose slope per voxel (2^-10) is far below the chamfer step:
    # every clamp is inactive, so the level is exactly half the pairwise average.
    planes = oracle.Planes(res, mips)
    x = np.arange(res, dtype=np.float64)
    fine = np.broadcast_to((x - 16) * 2.0 ** -10, (res, res, res)).astype(np.float16)
    planes.level(0, "dist")[:] = fine.view(np.uint16)
    planes.mat[:] = 0
    oracle.mip_refresh(planes, 1, band, 0.98, counts)
    pairs = (fine[..., 0::2].astype(np.float64) + fine[..., 1::2]) / 2
    want = (pairs[0::2, 0::2, :] * 0.5).astype(np.float16)
    assert np.array_equal(planes.level(1, "dist"), want.view(np.uint16))


# ---- SdfEditor apply (modify.ts:130-198) ------------------------------------------
def test_edit_flush_applies_a_sphere_cut_like_numpy(oracle):
    res, mips, world, origin, band = 64, 3, 8.0, (-4.0, -2.0, -4.0), 4.0
    planes = baked_planes(oracle, res, mips, world, origin, band)
    d0 = planes.level(0, "dist").copy()
    m0 = planes.level(0, "mat").copy()
    edit = sphere_edit((0.3, -0.5, 0.2), 0.6, smooth=0.05)
    counts = oracle.edit_flush(planes, world, origin, band, [edit])
    # numpy fold of one cut sphere, voxel by voxel, in f32
    voxel = world / res
    tv, bw = f32(voxel), f32(band * voxel)
    c = np.arange(res, dtype=f32)
    zz, yy, xx = np.meshgrid(c, c, c, indexing="ij")
    p = [f32(origin[0]) + (xx + f32(0.5)) * tv, f32(origin[1]) + (yy + f32(0.5)) * tv,
         f32(origin[2]) + (zz + f32(0.5)) * tv]
    v = [p[a] - f32(edit.pos[a]) for a in range(3)]
    dist = np.sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]) - f32(edit.size[0])
    acc0 = d0.view(np.float16).astype(f32) * bw
    acc1 = m0.view(np.float16).astype(f32)
    kk = np.maximum(f32(edit.blend), f32(1e-6))
    a = -dist
    h = np.clip(f32(0.5) - f32(0.5) * (acc0 - a) / kk, f32(0), f32(1))
    rx = (acc0 * (f32(1) - h) + a * h) + kk * h * (f32(1) - h)
    ny = acc1 * (f32(1) - h * f32(0.85)) + f32(edit.material) * (h * f32(0.85))
    new0 = acc0 * (f32(1) - f32(1)) + rx * f32(1)
    new1 = acc1 * (f32(1) - f32(1)) + ny * f32(1)
    want_d = np.clip(new0 / bw, f32(-1), f32(1)).astype(np.float16).view(np.uint16)
    want_m = new1.astype(np.float16).view(np.uint16)
    t0 = res // 8
    live0 = counts[:t0 ** 3].reshape((t0,) * 3) != 0
    vox_live = np.repeat(np.repeat(np.repeat(live0, 8, 0), 8, 1), 8, 2)
    got_d, got_m = planes.level(0, "dist"), planes.level(0, "mat")
    assert vox_live.any()
    assert np.array_equal(got_d[vox_live], want_d[vox_live])
    assert np.array_equal(got_m[vox_live], want_m[vox_live])
    assert np.array_equal(got_d[~vox_live], d0[~vox_live])
    # the carve reaches the surface: some voxel changed sign
    assert (d0.view(np.float16) < 0).sum() > (got_d.view(np.float16) < 0).sum()


def test_contacts_to_edits_match_brushset_make(oracle):
    pts = np.array([[0.1, 0.2, 0.3, 5.0], [1.0, -1.0, 0.5, 0.5], [-0.7, 0.0, 2.0, 3.5]], dtype=np.float32)
    edits = oracle.contacts_to_edits(3, pts, min_speed=1.0, radius=0.15, blend=0.03)
    assert len(edits) == 2
    for e, row in zip(edits, pts[[0, 2]]):
        want = make_brush({"kind": "sphere", "size": 0.15, "pos": tuple(float(v) for v in row[:3]), "op": "cut",
                           "smooth": 0.03})
        assert bytes(e) == bytes(want)
    # the atomic counter may exceed the list: only the first 32 points exist
    many = np.tile(pts[:1], (40, 1))
    assert len(oracle.contacts_to_edits(40, many, 1.0, 0.15, 0.03)) == 32


# ---- body centroids (pbd.ts:240-266) ---------------------------------------------
def test_body_centroids_follow_the_reference_reduction(oracle):
    rng = np.random.default_rng(5)
    pos = (rng.normal(size=(10000, 3)) * 3 + 7).astype(np.float32)
    rest = (rng.normal(size=(10000, 3))).astype(np.float32)
    bodies = [(0, 4096), (4096, 300), (4396, 5604), (9999, 0)]
    got = oracle.body_centroids(bodies, pos, rest, precision="f32")
    for b, (first, count) in enumerate(bodies):
        for q, src in ((0, pos), (3, rest)):
            lanes = np.zeros((256, 3), dtype=np.float32)
            for li in range(256):
                acc = np.zeros(3, dtype=np.float32)
                for t in range(li, count, 256):
                    acc = acc + src[first + t]
                lanes[li] = acc
            s = 128
            while s > 0:
                lanes[:s] = lanes[:s] + lanes[s:2 * s]
                s //= 2
            inv = np.float32(1) / np.float32(max(count, 1))
            assert np.array_equal(got[b, q:q + 3].astype(np.float32), lanes[0] * inv), (b, q)
    f64 = oracle.body_centroids(bodies, pos, rest, precision="f64")
    for b, (first, count) in enumerate(bodies[:3]):
        assert np.allclose(f64[b, :3], pos[first:first + count].astype(np.float64).mean(axis=0), rtol=1e-14)
        assert np.allclose(f64[b, 3:], rest[first:first + count].astype(np.float64).mean(axis=0), rtol=1e-12, atol=1e-15)
    assert np.all(f64[3] == 0)
