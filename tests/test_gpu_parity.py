"""HIP path vs the CPU oracle, bit for bit, on the 16k dam-break workload and on
small edge cases.  Every comparison here is exact: the plain kernels evaluate the
reference's f32 expressions in IEEE order with no contraction, as the oracle's
"WGSL-f32" mode does."""
import numpy as np
import pytest

import clayhip
import oracle as O
from clayhip import scenes
from gpu_helpers import bits_equal, describe_diff, make_pair, make_world, surface_opts

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small():
    return scenes.dam_break_small()


def test_keys_and_permutation_bit_exact(root, small):
    sim, orc, _ = make_pair(root, small)
    sim.substep()
    orc.substep()
    perm = sim.debug("perm")
    skeys = sim.debug("keys")
    k = orc.k
    G = k.grid_res ** 3
    want_keys = np.where(orc.keys == O.INVALID_KEY, G, orc.keys).astype(np.uint32)
    assert np.array_equal(perm, orc.perm), "stable (key, id) permutation differs"
    assert np.array_equal(skeys, want_keys[orc.perm]), "sorted keys differ"
    assert np.all(np.diff(skeys.astype(np.int64)) >= 0)
    sim.destroy()


def test_one_substep_bit_exact(root, small):
    sim, orc, _ = make_pair(root, small)
    sim.substep()
    orc.substep()
    for what in ("density", "pressure", "nvel", "pos", "vel"):
        got, want = sim.get_state(what), orc.get(what)
        assert bits_equal(got, want), describe_diff(what, got, want)
    sim.destroy()


@pytest.mark.parametrize("mode", ["plain", "tiled"])
def test_step_and_bake_bit_exact(root, small, mode):
    # Tiled: at mips 2-3 a brick spans 12-48 world units and holds all 16k particles,
    # so its list is cut into several chunks merged with atomicMin.
    sim, orc, _ = make_pair(root, small, splat_mode=mode)
    sim.step()
    for _ in range(small["options"]["substeps"]):
        orc.substep()
    got, want = sim.get_state("pos"), orc.get("pos")
    assert bits_equal(got, want), describe_diff("pos", got, want)
    o = surface_opts(small)
    grids = O.splat_bake(o, orc.get("pos"))
    tex = O.splat_resolve(o, grids)
    for m in range(o.mips):
        codes = sim.surface.codes(m)
        assert np.array_equal(codes, grids[m]), f"mip {m}: {(codes != grids[m]).sum()} codes differ"
        assert np.array_equal(sim.surface.volume.download(m), tex[m]), f"mip {m}: texels differ"
    got_pts = sim.readContacts()
    raw, want_pts = orc.read_contacts()
    assert sim.lastContactCount == raw
    assert got_pts == [tuple(float(np.float32(v)) for v in p) for p in want_pts]
    sim.destroy()


def test_surface_codes_while_the_column_collapses(root, small):
    """40 frames of the 16k dam break with the tiled bake: the column falls, hits
    the floor and splashes, so the bake's interior radii (voxels beyond them are
    skipped) meet ordered, compressed and scattered neighbourhoods.  Codes, texels
    and live tiles must equal the oracle's full splat of the same positions."""
    opts = dict(small["options"], splatMode="tiled")
    world, _, vol = make_world(root, small)
    sim = clayhip.FluidSim(root, world, opts)
    sim.set_state(small["pos"])
    o = surface_opts(small)
    for frame in range(1, 41):
        sim.step()
        if frame not in (10, 25, 40):
            continue
        grids = O.splat_bake(o, sim.get_state("pos"))
        tex = O.splat_resolve(o, grids)
        for m in range(o.mips):
            codes = sim.surface.codes(m)
            assert np.array_equal(codes, grids[m]), f"frame {frame} mip {m}: {(codes != grids[m]).sum()} codes differ"
            assert np.array_equal(sim.surface.volume.download(m), tex[m]), f"frame {frame} mip {m}: texels differ"
    sim.destroy()
    world.destroy()
    if vol is not None:
        vol.destroy()


def test_free_run_ten_steps_bit_exact(root, small):
    sim, orc, _ = make_pair(root, small)
    for step in range(10):
        sim.step()
        for _ in range(small["options"]["substeps"]):
            orc.substep()
        got, want = sim.get_state("pos"), orc.get("pos")
        assert bits_equal(got, want), f"step {step}: " + describe_diff("pos", got, want)
    assert bits_equal(sim.get_state("vel"), orc.get("vel"))
    sim.destroy()


@pytest.mark.parametrize("layers", [1, 3])
def test_contacts_match_oracle_in_id_order(root, layers):
    """The list keeps the first 32 contacts by particle id (the order the oracle
    appends in); the raw count is the reference's counter.  One layer gives fewer
    than 32 contacts, three layers more."""
    pos = scenes.lattice(12, layers, 12, 0.05, (-0.3, -0.42, -0.3), seed=3)
    if layers == 1:
        pos = pos[:20]
    scene = scenes.dam_break_small()
    scene = dict(scene, pos=pos, options=dict(scene["options"], capacity=pos.shape[0]))
    sim, orc, _ = make_pair(root, scene)
    sim.substep()
    orc.substep()
    got = sim.readContacts()
    raw, want = orc.read_contacts()
    assert sim.lastContactCount == raw and raw > 0
    assert (raw <= 32) == (layers == 1)
    assert got == [tuple(float(np.float32(v)) for v in p) for p in want]
    sim.destroy()


def test_bucket_overflow_matches_oracle(root):
    # 200 particles squeezed into one hash cell: the bucket keeps the first 64 ids.
    rng = np.random.default_rng(5)
    pos = (rng.random((200, 3), dtype=np.float32) * np.float32(0.04) + np.float32([0.31, 0.51, 0.71]))
    scene = scenes.dam_break_small()
    scene = dict(scene, pos=pos.astype(np.float32), options=dict(scene["options"], capacity=256))
    sim, orc, _ = make_pair(root, scene)
    sim.substep()
    orc.substep()
    for what in ("density", "nvel", "pos"):
        got, want = sim.get_state(what), orc.get(what)
        assert bits_equal(got, want), describe_diff(what, got, want)
    sim.destroy()


def test_particles_outside_grid_match_oracle(root):
    # Some particles outside the hash grid (not inserted) and some on the walls.
    rng = np.random.default_rng(9)
    pos = rng.uniform(-14, 22, size=(2000, 3)).astype(np.float32)
    scene = scenes.dam_break_small()
    scene = dict(scene, pos=pos, options=dict(scene["options"], capacity=2000))
    sim, orc, _ = make_pair(root, scene)
    sim.substep()
    orc.substep()
    for what in ("density", "nvel", "pos", "vel"):
        got, want = sim.get_state(what), orc.get(what)
        assert bits_equal(got, want), describe_diff(what, got, want)
    sim.destroy()


def test_field_eval_bit_exact(root):
    brushes = scenes.ridge_world_brushes()
    recs = [clayhip.make_brush(b) for b in brushes]
    pts = np.random.default_rng(2).uniform(-12, 12, size=(4096, 3)).astype(np.float32)
    pts[:, 1] = np.random.default_rng(3).uniform(-3, 3, size=4096).astype(np.float32)
    f = clayhip.analyticField(root, brushes)
    d, n = f.evaluate(pts)
    od, on = O.Collider().add_analytic(recs).eval(pts)
    assert bits_equal(d, od), describe_diff("analytic dist", d, od)
    assert bits_equal(n, on), describe_diff("analytic normal", n, on)
    vol = clayhip.SdfVolume(root, resolution=64, mipLevels=3, worldSize=24, origin=(-12, -4, -12), band=4)
    vol.bakeBrushes(brushes)
    fv = clayhip.volumeField(vol)
    d, n = fv.evaluate(pts)
    od, on = O.Collider().add_volume(vol.download(0)[..., 0], 64, (-12, -4, -12), 24, 4).eval(pts)
    assert bits_equal(d, od), describe_diff("volume dist", d, od)
    assert bits_equal(n, on), describe_diff("volume normal", n, on)
    u = clayhip.unionField(f, fv)
    d, n = u.evaluate(pts)
    ou = O.Collider().add_analytic(recs).add_volume(vol.download(0)[..., 0], 64, (-12, -4, -12), 24, 4)
    od, on = ou.eval(pts)
    assert bits_equal(d, od) and bits_equal(n, on)
    for x in (u, fv, f):
        x.destroy()
    vol.destroy()


def test_brush_bake_bit_exact(root):
    brushes = scenes.ridge_world_brushes()
    recs = [clayhip.make_brush(b) for b in brushes]
    vol = clayhip.SdfVolume(root, resolution=64, mipLevels=3, worldSize=24, origin=(-12, -4, -12), band=4)
    vol.bakeBrushes(brushes)
    want = O.bake_brushes(64, 3, 24.0, (-12, -4, -12), 4.0, recs)
    for m in range(3):
        assert np.array_equal(vol.download(m), want[m]), f"mip {m} differs"
    vol.destroy()


def _cloud(pos: np.ndarray):
    import torch

    t = torch.from_numpy(np.ascontiguousarray(pos.T)).cuda()
    return t, clayhip.ParticleCloud(pos.shape[0], (t[0], t[1], t[2]), live=("prefix", pos.shape[0]))


@pytest.mark.parametrize("owner", [False, True], ids=["brick", "owner"])
@pytest.mark.parametrize("res,mips,band,radius", [(64, 4, 4, 0.05), (32, 3, 2, 0.3), (80, 2, 4, 0.12)])
def test_tiled_splat_equals_plain_and_oracle(root, res, mips, band, radius, owner):
    rng = np.random.default_rng(res)
    pos = rng.uniform(-2, 3, size=(6000, 3)).astype(np.float32)
    pos[:2000] = rng.uniform(-12.5, 20.5, size=(2000, 3))  # some outside the volume
    t, cloud = _cloud(pos)
    common = dict(radius=radius, resolution=res, mipLevels=mips, band=band, worldSize=24, origin=(-12, -4, -12))
    plain = clayhip.SplatField(root, cloud, common, mode="plain")
    tiled = clayhip.SplatField(root, cloud, common, mode="tiled", keepCodes=True)
    tiled.setSchedule(chunk=512 if owner else 2048, groups_per_cu=4, owner=owner)
    o = O.splat_opts(res, mips, 24.0, (-12, -4, -12), band, radius)
    want = O.splat_bake(o, pos)
    for rep in range(2):  # the second bake must clear the first bake's tiles
        if rep == 1:
            t[1] += 0.7
            want = O.splat_bake(o, t.cpu().numpy().T)
        plain.bake()
        tiled.bake()
        for m in range(mips):
            assert np.array_equal(plain.codes(m), want[m]), f"plain mip {m}"
            assert np.array_equal(tiled.codes(m), want[m]), f"tiled mip {m}"
            assert np.array_equal(tiled.volume.download(m), plain.volume.download(m)), f"texels mip {m}"
            assert np.array_equal(tiled.liveTiles(m), plain.liveTiles(m)), f"live tiles mip {m}"
    plain.destroy()
    tiled.destroy()


@pytest.mark.parametrize("radius", [0.2, 0.05])
def test_splat_with_normals_and_range_rule(root, radius):
    """radius 0.2 spans more than 5 voxels here, so the owner schedule falls back to
    bricks; at 0.05 it runs."""
    import torch

    rng = np.random.default_rng(11)
    pos = rng.uniform(-1, 1, size=(3000, 3)).astype(np.float32)
    nrm = rng.normal(size=(3000, 3)).astype(np.float32)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    tp = torch.from_numpy(np.ascontiguousarray(pos.T)).cuda()
    tn = torch.from_numpy(np.ascontiguousarray(nrm.T)).cuda()
    cloud = clayhip.ParticleCloud(3000, (tp[0], tp[1], tp[2]), live=("range", 500, 1500),
                                  normals=(tn[0], tn[1], tn[2]))
    opts = dict(radius=radius, resolution=32, mipLevels=2, band=4, worldSize=2.7, origin=(-1.35, -1.35, -1.35))
    o = O.splat_opts(32, 2, 2.7, (-1.35, -1.35, -1.35), 4, radius)
    want = O.splat_bake(o, pos, 500, 1500, normals=nrm)
    for mode in ("plain", "tiled", "owner"):
        s = clayhip.SplatField(root, cloud, opts, mode="plain" if mode == "plain" else "tiled", keepCodes=True)
        if mode == "owner":
            s.setSchedule(chunk=256, owner=True)
        s.bake()
        for m in range(2):
            assert np.array_equal(s.codes(m), want[m]), f"{mode} mip {m}"
        s.destroy()


def test_set_center_follows_and_far_origin_saturates(root):
    pos = scenes.lattice(8, 8, 8, 0.1, (0, 0, 0), seed=2)
    t, cloud = _cloud(pos)
    opts = dict(radius=0.15, resolution=32, mipLevels=3, band=4, worldSize=2.7, origin=(-1, -1, -1))
    for mode in ("plain", "tiled", "owner"):
        s = clayhip.SplatField(root, cloud, opts, mode="plain" if mode == "plain" else "tiled", keepCodes=True)
        if mode == "owner":
            s.setSchedule(owner=True)
        s.setCenter((0.4, 0.4, 0.4))
        s.bake()
        o = O.splat_opts(32, 3, 2.7, s.volume.origin, 4, 0.15)
        want = O.splat_bake(o, pos)
        for m in range(3):
            assert np.array_equal(s.codes(m), want[m])
        # Far away: every voxel reads saturated +1 (splat.ts:289-295).
        s.setCenter((100.0, 100.0, 100.0))
        s.bake()
        for m in range(3):
            assert np.all(s.codes(m) == 65535)
            assert np.all(s.volume.download(m)[..., 0] == 0x3C00)
            assert not s.liveTiles(m).any()
        s.destroy()


def test_spawn_emits_into_the_ring_like_the_oracle(root, small):
    """FluidSim.spawn + step (fluid.ts:268-291, 511-549): emission slots, live count,
    ring wrap.  Emission uses pow(x, 1/3), whose last bit is the math library's
    (parity unpinned), so positions are compared to 1e-5 after the step's substeps."""
    opts = dict(small["options"], capacity=1500)
    brushes = small["collider"]["brushes"]
    world = clayhip.analyticField(root, brushes)
    sim = clayhip.FluidSim(root, world, opts)
    orc = O.Fluid(opts, None, collider=O.Collider().add_analytic([clayhip.make_brush(b) for b in brushes]),
                  capacity=1500)
    for rep, (pos, vel) in enumerate([((0.0, 2.0, 0.0), (0.0, -1.0, 0.0)), ((0.3, 2.5, -0.2), (1.0, 0.0, 0.0))]):
        sim.spawn(1000, pos, vel)
        orc.spawn(1000, pos, vel)
        sim.step()
        orc.emit_pending()
        for _ in range(opts["substeps"]):
            orc.substep()
        assert sim.particleCount == orc.n == (1000 if rep == 0 else 1500)
        got, want = sim.get_state("pos"), orc.get("pos")
        assert np.abs(got - want).max() < 1e-5, describe_diff("pos", got, want)
    sim.destroy()
    world.destroy()


def test_empty_fluid_bakes_saturated(root, small):
    world = clayhip.analyticField(root, small["collider"]["brushes"])
    sim = clayhip.FluidSim(root, world, dict(small["options"], splatMode="tiled"))
    sim.step()
    assert sim.particleCount == 0
    assert np.all(sim.surface.volume.download(0)[..., 0] == 0x3C00)
    sim.destroy()
    world.destroy()
