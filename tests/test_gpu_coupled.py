"""The coupled frame (fluid + ClaySolver body + eroding world) over batched scenes.

Each object on its own is checked against the oracle elsewhere.  Here: scene s of a
3-scene coupled batch must equal, bit for bit, the same scene run as a 1-scene
coupled batch (fluid state, body particles, the world instance after erosion at
every mip, the body volume), which checks every per-scene offset of the coupling; the body volume
must centre on the FP64 centroid; erosion must touch only binned tiles.
"""
import numpy as np
import pytest

import oracle as O
from clayhip import scenes
from clayhip.coupled import CoupledBatch

pytestmark = pytest.mark.gpu

WORLD = {"resolution": 64, "mipLevels": 3, "worldSize": 8.0, "origin": (-4.0, -2.0, -4.0), "band": 4}
BRUSHES = [{"kind": "roundBox", "size": (3.5, 0.5, 3.5), "radius": 0.25, "pos": (0, -1.0, 0)},
           {"kind": "sphere", "size": 0.8, "pos": (0.9, -0.4, 0.3), "op": "cut"}]
PHASES = [0.0, 2.0, 4.0]
SEEDS = [21, 22, 23]


def blocks():
    out = [scenes.drop_block(c, d, seed=40 + i) for i, (c, d) in enumerate([(2400, 0.3), (1800, 0.6), (3000, 0.2)])]
    return [np.asarray(b) for b in out]


def build(root, which):
    bl = blocks()
    counts = [bl[i].shape[0] for i in which]
    pos = np.concatenate([bl[i] for i in which], axis=0)
    opts = dict(scenes.batch_options(), precision="f32")
    return CoupledBatch(root, counts, pos, opts, WORLD, BRUSHES, body_count=512, body_radius=0.4,
                        body_box=1.2, body_res=32, body_mips=2, body_splat_radius=0.03,
                        erosion=(0.3, 0.12, 0.03), orbit=0.8, height=0.2,
                        phases=[PHASES[i] for i in which], body_seeds=[SEEDS[i] for i in which])


def test_batched_coupling_equals_single_scene_coupling(root):
    import torch

    frames = 3
    batch = build(root, [0, 1, 2])
    stream = torch.cuda.current_stream()
    for _ in range(frames):
        batch.step(stream)
    torch.cuda.synchronize()
    bp = batch.fluid.get_state("pos")
    bb = batch.body_positions()
    eroded = [batch.editor.readTileCounts(s) for s in range(3)]
    assert sum(e[0] for e in eroded) > 0, "no erosion happened"
    for s in range(3):
        single = build(root, [s])
        for _ in range(frames):
            single.step(stream)
        torch.cuda.synchronize()
        first, n = batch.fluid.scene(s)
        sp = single.fluid.get_state("pos")
        assert np.array_equal(bp[first:first + n].view(np.uint32), sp.view(np.uint32)), f"scene {s} fluid"
        assert np.array_equal(bb[s * 512:(s + 1) * 512].view(np.uint32), single.body_positions().view(np.uint32)), \
            f"scene {s} body particles"
        for m in range(3):
            assert np.array_equal(batch.world.download(m, instance=s), single.world.download(m)), f"scene {s} world mip {m}"
        for m in range(2):
            assert np.array_equal(batch.body.sceneCodes(s, m), single.body.sceneCodes(0, m)), f"scene {s} body mip {m}"
        assert batch.editor.readTileCounts(s) == single.editor.readTileCounts(0)
        single.destroy()
    batch.destroy()


def test_body_volume_follows_the_fp64_centroid(root):
    import torch

    c = build(root, [0, 1])
    stream = torch.cuda.current_stream()
    c.step(stream)
    torch.cuda.synchronize()
    pos = c.body_positions()
    fc = [(s * 512, 512) for s in range(2)]
    com = O.body_centroids(fc, pos, precision="f64")[:, :3]
    assert np.array_equal(c.body_com.cpu().numpy().reshape(2, 6)[:, :3].view(np.uint64), com.view(np.uint64))
    half = 1.2 / 2
    for s in range(2):
        want = (com[s] - half).astype(np.float32)
        o = O.splat_opts(32, 2, 1.2, [float(v) for v in want], 4, 0.03)
        grids = O.splat_bake(o, pos[s * 512:(s + 1) * 512])
        for m in range(2):
            assert np.array_equal(c.body.sceneCodes(s, m), grids[m]), f"scene {s} body mip {m}"
    c.destroy()


def test_erosion_touches_only_binned_tiles(root):
    import torch

    c = build(root, [2])
    before = [c.world.download(m) for m in range(3)]
    stream = torch.cuda.current_stream()
    c.step(stream)
    torch.cuda.synchronize()
    counts = c.editor.cellCounts()
    _, tile_res, offs, _ = O.volume_derive(64, 3)
    for m in range(3):
        tr = tile_res[m]
        live = counts[offs[m]:offs[m] + tr ** 3].reshape((tr,) * 3) != 0
        vox = np.repeat(np.repeat(np.repeat(live, 8, 0), 8, 1), 8, 2)
        after = c.world.download(m)
        assert np.array_equal(after[~vox], before[m][~vox]), f"mip {m}: a tile outside the edit mask changed"
    c.destroy()
