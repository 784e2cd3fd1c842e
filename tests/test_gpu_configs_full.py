"""The three large workloads at their full sizes, against the oracle.

  * 1M dam break: 1 048 576 particles against the 288^3 x 3 baked ridge world,
    surface baked at 288^3 x 3 mips.  One step: positions and every mip's codes
    bit for bit.
  * Batched: the 256 scenes (3 532 015 particles, counts log-uniform in
    [1024, 65536]) with the tuned launch shapes.  One f32 frame: every scene's
    positions and its following volume's codes equal the oracle's step of that
    scene alone, bit for bit; the FP16-LDS frame is within the precision
    variant's tolerance of the f32 one.
  * Coupled: 64 scenes x (524 288 fluid + 4 096 body particles) over 288^3 x 3
    worlds.  One frame: for two scenes every stage against the oracle (body solver
    step, FP64 centroids, body bake at the centroid, fluid step against
    union(world, body volume), the fluid surface's codes, contacts, and the eroded
    world at every mip after the flush); one scene against its own solo run; FP64
    mass of every scene; FP16-LDS within tolerance of f32.
"""
import math

import numpy as np
import pytest

import clayhip
import oracle as O
from clayhip import autotune, make_brush, scenes
from clayhip.coupled import FRAME, SUBSTEPS, CoupledBatch, coupled_scene
from gpu_helpers import bits_equal, describe_diff

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]


def tuned(workload: str, precision: str) -> dict | None:
    return autotune.load(autotune.DEFAULT_CACHE, autotune.cache_key(workload, precision))


def print_case(msg: str) -> None:
    print(msg, flush=True)


# ---- 1M single scene --------------------------------------------------------------
def test_dam_break_1m_step_against_the_oracle(root, oracle):
    sc = scenes.dam_break_large()
    print_case("1M dam break, 288^3 world, 288^3 x 3 bake")
    v = sc["collider"]["volume"]
    recs = [make_brush(b) for b in sc["collider"]["brushes"]]
    vol = clayhip.SdfVolume(root, v)
    vol.bakeBrushes(sc["collider"]["brushes"])
    want_tex = O.bake_brushes(v["resolution"], v["mipLevels"], v["worldSize"], v["origin"], v["band"], recs)
    for m in range(v["mipLevels"]):
        assert np.array_equal(vol.download(m), want_tex[m]), f"world mip {m}"
    world = clayhip.volumeField(vol)
    opts = dict(sc["options"], splatMode="tiled")
    sim = clayhip.FluidSim(root, world, opts)
    t = tuned("dam-break-1m", "f32")
    if t:
        sim.tuning = t
    sim.set_state(sc["pos"])
    sim.step()
    got = sim.get_state("pos")
    col = O.Collider().add_volume(want_tex[0][..., 0], v["resolution"], v["origin"], v["worldSize"], v["band"])
    orc = O.Fluid(opts, sc["pos"], collider=col)
    for _ in range(opts["substeps"]):
        orc.substep()
    want = orc.get("pos")
    assert bits_equal(got, want), describe_diff("positions", got, want)
    raw, pts = orc.read_contacts()
    got_pts = sim.readContacts()
    assert sim.lastContactCount == raw and sorted(got_pts) == sorted(pts)
    so = O.splat_opts(opts["bakeResolution"], opts["bakeMipLevels"], opts["worldSize"], opts["origin"],
                      opts["bakeBand"], opts["surfaceRadius"])
    grids = O.splat_bake(so, want)
    for m in range(opts["bakeMipLevels"]):
        c = sim.surface.codes(m)
        assert np.array_equal(c, grids[m]), f"codes of mip {m}: {(c != grids[m]).sum()} differ"
    live0 = int(sim.surface.liveTiles(0).sum())
    assert 0 < live0 < (288 // 8) ** 3
    sim.destroy()
    world.destroy()
    vol.destroy()


# ---- batched 256 ------------------------------------------------------------------
def oracle_scene_step(opts, pos, col):
    """One step of one scene in the oracle: its substeps, then the splat of its
    following volume centred on the f32 midpoint of its bounds."""
    fl = O.Fluid(dict(opts, capacity=pos.shape[0]), pos, collider=col)
    for _ in range(opts["substeps"]):
        fl.substep()
    p = fl.get("pos")
    centre = (p.min(axis=0) + p.max(axis=0)) * np.float32(0.5)
    origin = [float(np.float32(c - opts["followSize"] / 2)) for c in centre.astype(np.float64)]
    so = O.splat_opts(opts["bakeResolution"], opts["bakeMipLevels"], opts["followSize"], origin, opts["bakeBand"],
                      opts["surfaceRadius"])
    return p, origin, O.splat_bake(so, p)


def test_batched_256_frame_against_the_oracle(root, oracle):
    sc = scenes.batched(256)
    print_case(f"batched 256 scenes, {int(sc['counts'].sum())} particles")
    world = clayhip.analyticField(root, sc["collider"]["brushes"])
    col = O.Collider().add_analytic([make_brush(b) for b in sc["collider"]["brushes"]])
    shapes = tuned("batched-256", "fp16lds")
    batches = {}
    for precision in ("f32", "fp16lds"):
        b = clayhip.FluidBatch(root, world, dict(sc["options"], precision=precision), sc["counts"])
        if shapes:
            b.tuning = dict(shapes, exact_lds=1)
        b.set_state(sc["pos"])
        b.step()
        batches[precision] = b
    b = batches["f32"]
    pos = b.get_state("pos")
    opts = sc["options"]
    for s in range(b.scenes):
        first, n = b.scene(s)
        want, origin, grids = oracle_scene_step(opts, sc["pos"][first:first + n], col)
        assert bits_equal(pos[first:first + n], want), describe_diff(f"scene {s} positions", pos[first:first + n], want)
        assert np.array_equal(b.origin(s).view(np.uint32), np.float32(origin).view(np.uint32)), f"scene {s} origin"
        for m in range(b.mips):
            assert np.array_equal(b.codes(s, m), grids[m].reshape(-1)), f"scene {s} codes of mip {m}"
    h = b.constants.h
    d = np.linalg.norm(batches["fp16lds"].get_state("pos").astype(np.float64) - pos, axis=1) / h
    assert np.sqrt(np.mean(d * d)) < 1e-4 and d.max() < 5e-2, (np.sqrt(np.mean(d * d)), d.max())
    for x in batches.values():
        x.destroy()
    world.destroy()


# ---- coupled, 64 scenes -------------------------------------------------------------
BODY = dict(count=4096, radius=0.95, box=2.7, res=32, mips=2, splat_radius=0.08)
EROSION = (2.0, 0.12, 0.03)
CHECK = (0, 37)


def body_spacing():
    return BODY["radius"] * (4.0 / 3.0 * math.pi / BODY["count"]) ** (1.0 / 3.0)


def test_coupled_64_frame_against_the_oracle(root, oracle):
    import torch

    sc = coupled_scene(scenes=64)
    S = 64
    print_case(f"coupled 64 scenes x ({int(sc['counts'][0])} + 4096), 288^3 worlds")
    w = sc["world"]
    recs = [make_brush(b) for b in sc["world_brushes"]]
    world_tex = O.bake_brushes(w["resolution"], w["mipLevels"], w["worldSize"], w["origin"], w["band"], recs)
    stream = torch.cuda.current_stream()
    runs = {}
    for precision in ("f32", "fp16lds"):
        c = CoupledBatch(root, sc["counts"], sc["pos"], dict(sc["options"], precision=precision), w,
                         sc["world_brushes"])
        if precision == "f32":
            for m in range(w["mipLevels"]):
                assert np.array_equal(c.world.download(m, instance=S - 1), world_tex[m]), f"world mip {m}"
            body0 = c.body_positions()
            rest = c.solver.get("rest")
        c.step(stream)
        torch.cuda.synchronize()
        runs[precision] = c
    c = runs["f32"]
    # FP64 centroids of every body, from the body particles the solver left
    bpos = c.body_positions()
    fc = [(s * BODY["count"], BODY["count"]) for s in range(S)]
    com = O.body_centroids(fc, bpos, precision="f64")[:, :3]
    got_com = c.body_com.cpu().numpy().reshape(S, 6)[:, :3]
    assert np.array_equal(got_com.view(np.uint64), com.view(np.uint64)), "body centroids"
    fpos = c.fluid.get_state("pos")
    opts = sc["options"]
    spacing = body_spacing()
    plasticity = 1.0 - (1.0 - 0.35) ** (FRAME / SUBSTEPS)
    for s in CHECK:
        print_case(f"coupled scene {s} against the oracle")
        world_col = O.Collider().add_volume(world_tex[0][..., 0], w["resolution"], w["origin"], w["worldSize"],
                                            w["band"])
        # 1. the body's solver step against the scene's world
        sl = slice(s * BODY["count"], (s + 1) * BODY["count"])
        so = O.Solver(BODY["count"], 1, gravity=(0.0, -18.0, 0.0), dt=FRAME, substeps=SUBSTEPS,
                      radius=0.6 * spacing, friction=0.45, damping=0.997, max_speed=26.0, collider=world_col)
        so.set_bodies([dict(first=0, count=BODY["count"], stiffness=0.55, plasticity=plasticity)])
        so.set_particles(body0[sl], rest[sl], np.zeros(BODY["count"], np.uint32))
        so.step()
        assert bits_equal(bpos[sl], so.get(0)), describe_diff(f"scene {s} body", bpos[sl], so.get(0))
        # 2. the body's volume at its FP64 centroid
        borigin = [float(np.float32(v - BODY["box"] / 2)) for v in com[s]]
        bo = O.splat_opts(BODY["res"], BODY["mips"], BODY["box"], borigin, 4, BODY["splat_radius"])
        bgrids = O.splat_bake(bo, bpos[sl])
        for m in range(BODY["mips"]):
            assert np.array_equal(c.body.sceneCodes(s, m), bgrids[m]), f"scene {s} body codes of mip {m}"
        btex = O.splat_resolve(bo, bgrids)
        assert np.array_equal(c.body.volume.download(0, instance=s)[..., 0], btex[0][..., 0]), f"scene {s} body texels"
        # 3. the fluid step against union(world, body volume)
        first, n = c.fluid.scene(s)
        col = O.Collider().add_volume(world_tex[0][..., 0], w["resolution"], w["origin"], w["worldSize"],
                                      w["band"]).add_volume(btex[0][..., 0], BODY["res"], borigin, BODY["box"], 4)
        fl = O.Fluid(dict(opts, capacity=n), sc["pos"][first:first + n], collider=col)
        for _ in range(opts["substeps"]):
            fl.substep()
        want = fl.get("pos")
        assert bits_equal(fpos[first:first + n], want), describe_diff(f"scene {s} fluid", fpos[first:first + n], want)
        # 4. the fluid surface's codes at its following origin
        fo = c.fluid.origin(s)
        fso = O.splat_opts(opts["bakeResolution"], opts["bakeMipLevels"], opts["followSize"],
                           [float(v) for v in fo], opts["bakeBand"], opts["surfaceRadius"])
        fgrids = O.splat_bake(fso, want)
        for m in range(opts["bakeMipLevels"]):
            assert np.array_equal(c.fluid.codes(s, m), fgrids[m].reshape(-1)), f"scene {s} surface mip {m}"
        # 5. contacts, then the erosion flush of the world instance
        raw, pts = fl.read_contacts()
        graw, gpts = c.fluid.read_contacts(s)
        assert graw == raw and np.array_equal(np.float32(gpts), np.float32(pts)), f"scene {s} contacts"
        edits = O.contacts_to_edits(raw, np.array(pts, dtype=np.float32).reshape(-1, 4), *EROSION)
        planes = O.Planes(w["resolution"], w["mipLevels"], world_tex)
        counts = O.edit_flush(planes, w["worldSize"], w["origin"], w["band"], edits)
        assert np.array_equal(c.editor.cellCounts(s), counts), f"scene {s} edit tiles"
        for m in range(w["mipLevels"]):
            assert np.array_equal(c.world.download(m, instance=s), planes.texels(m)), f"scene {s} world mip {m}"
    # every scene: FP64 mass; FP16-LDS within tolerance of f32
    m = c.fluid.constants.mass
    for s in range(S):
        assert c.fluid.conserved(s)["mass"] == pytest.approx(m * int(sc["counts"][s]), rel=1e-15)
    # The bodies start inside the fluid blocks: the first frame pushes those particles
    # onto the body surfaces and piles them far above rest density, where the f16
    # offsets' density error is multiplied by the stiffness.  The bulk stays close.
    d = np.linalg.norm(runs["fp16lds"].fluid.get_state("pos").astype(np.float64) - fpos, axis=1) / 0.1
    stats = (float(np.sqrt(np.mean(d * d))), float(np.quantile(d, 0.99)), float(np.quantile(d, 0.999)), float(d.max()))
    print_case(f"coupled FP16-LDS vs f32 after one frame, /h: rms {stats[0]:.3g} p99 {stats[1]:.3g} "
               f"p99.9 {stats[2]:.3g} max {stats[3]:.3g}")
    assert stats[0] < 5e-4 and stats[1] < 2e-3, stats
    runs["fp16lds"].destroy()
    # one scene against its solo run
    s = 41
    print_case(f"coupled scene {s} against its solo run")
    first, n = c.fluid.scene(s)
    solo = CoupledBatch(root, sc["counts"][s:s + 1], sc["pos"][first:first + n], dict(opts, precision="f32"), w,
                        sc["world_brushes"], phases=[s * (2 * math.pi / S)], body_seeds=[11 + s])
    solo.step(stream)
    torch.cuda.synchronize()
    assert bits_equal(fpos[first:first + n], solo.fluid.get_state("pos")), f"scene {s} fluid vs solo"
    assert bits_equal(bpos[s * BODY["count"]:(s + 1) * BODY["count"]], solo.body_positions()), f"scene {s} body"
    for m in range(w["mipLevels"]):
        assert np.array_equal(c.world.download(m, instance=s), solo.world.download(m)), f"scene {s} world mip {m}"
    solo.destroy()
    c.destroy()
