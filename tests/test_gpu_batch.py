"""Batched scenes against single FluidSims, against eager frames, and against the oracle.

The batch runs every scene's FluidSim with the same launches: keys carry the
scene, each scene has its own cell records, contacts and following volume.  In
f32 mode each scene's particles, densities and splat codes must equal, bit for
bit, those of a FluidSim holding only that scene.  In FP16-LDS mode the staging
decision depends on which particles share a wave, so there the comparison is the
precision variant's tolerance.
"""
import numpy as np
import pytest

import clayhip
import oracle as O
from clayhip import scenes
from gpu_helpers import bits_equal, describe_diff

pytestmark = pytest.mark.gpu

COUNTS = [1500, 3001, 800, 5000, 1024, 2222]


def small_batch(seed=3):
    counts = np.array(COUNTS, dtype=np.uint32)
    rng_drops = [0.5, 2.9, 1.2, 0.7, 2.0, 1.6]
    blocks = [scenes.drop_block(int(c), d, seed=seed * 100 + i) for i, (c, d) in enumerate(zip(counts, rng_drops))]
    return counts, np.concatenate(blocks, axis=0)


def make_batch(root, world, opts, counts, pos, graph=True):
    b = clayhip.FluidBatch(root, world, opts, counts)
    b.set_graph(graph)
    b.set_state(pos)
    return b


@pytest.fixture(scope="module")
def floor(root):
    w = clayhip.analyticField(root, [scenes.floor_brush()])
    yield w
    w.destroy()


def test_batch_equals_single_fluidsims(root, floor):
    counts, pos = small_batch()
    opts = scenes.batch_options()
    b = make_batch(root, floor, opts, counts, pos)
    steps = 3
    for _ in range(steps):
        b.step()
    bp, bv, brho = b.get_state("pos"), b.get_state("vel"), b.get_state("density")
    for s in range(len(counts)):
        first, n = b.scene(s)
        sim = clayhip.FluidSim(root, floor, dict(opts, capacity=n))
        sim.set_state(pos[first:first + n])
        for _ in range(steps):
            sim.step()
        sl = slice(first, first + n)
        assert bits_equal(bp[sl], sim.get_state("pos")), describe_diff(f"scene {s} pos", bp[sl], sim.get_state("pos"))
        assert bits_equal(bv[sl], sim.get_state("vel")), f"scene {s} vel"
        assert bits_equal(brho[sl], sim.get_state("density")), f"scene {s} density"
        for m in range(b.mips):
            assert np.array_equal(b.codes(s, m), sim.surface.codes(m).reshape(-1)), f"scene {s} mip {m} codes"
            assert np.array_equal(b.live_tiles(s, m), sim.surface.liveTiles(m).reshape(-1)), f"scene {s} mip {m} tiles"
            assert np.array_equal(b.download(s, m), sim.surface.volume.download(m)), f"scene {s} mip {m} texels"
        raw_b, pts_b = b.read_contacts(s)
        pts_s = sim.readContacts()
        assert raw_b == sim.lastContactCount
        if raw_b <= clayhip._lib.CONTACTS:
            assert sorted(pts_b) == sorted(pts_s)
        sim.destroy()
    b.destroy()


def test_graph_replays_equal_eager_frames(root, floor):
    counts, pos = small_batch(seed=5)
    opts = scenes.batch_options()
    g = make_batch(root, floor, opts, counts, pos, graph=True)
    e = make_batch(root, floor, opts, counts, pos, graph=False)
    for frame in range(4):
        g.step()
        e.step()
        # host work and synchronisation between replays
        gp, ep = g.get_state("pos"), e.get_state("pos")
        assert bits_equal(gp, ep), describe_diff(f"frame {frame} pos", gp, ep)
        for s in (0, 3):
            for m in range(g.mips):
                assert np.array_equal(g.codes(s, m), e.codes(s, m)), f"frame {frame} scene {s} mip {m}"
    g.destroy()
    e.destroy()


def test_following_origin_and_codes_match_the_oracle(root, floor):
    counts, pos = small_batch(seed=9)
    opts = scenes.batch_options()
    b = make_batch(root, floor, opts, counts, pos)
    b.step()
    p = b.get_state("pos")
    half = opts["followSize"] / 2
    for s in range(len(counts)):
        first, n = b.scene(s)
        q = p[first:first + n]
        centre = (q.min(axis=0) + q.max(axis=0)) * np.float32(0.5)
        want = (centre.astype(np.float64) - half).astype(np.float32)
        got = b.origin(s)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (s, got, want)
    # codes of two scenes against the oracle's splat at the device origin
    for s in (1, 4):
        first, n = b.scene(s)
        o = O.splat_opts(opts["bakeResolution"], opts["bakeMipLevels"], opts["followSize"],
                         [float(v) for v in b.origin(s)], opts["bakeBand"], opts["surfaceRadius"])
        grids = O.splat_bake(o, p[first:first + n])
        for m in range(b.mips):
            assert np.array_equal(b.codes(s, m), grids[m].reshape(-1)), f"scene {s} mip {m}"
    b.destroy()


def test_codes_after_impact_match_the_oracle(root, floor):
    """The surface bake skips, for each particle surrounded by neighbours in all 26
    lattice directions, the voxels beyond its interior radius.  After 30 frames the
    low blocks have hit the floor and spread (irregular neighbourhoods, many
    particles with and without full cover); every scene's codes must still be the
    oracle's full splat of the same positions, for the FP16-LDS frame as for f32."""
    counts, pos = small_batch(seed=5)
    for precision in ("fp16lds", "f32"):
        opts = dict(scenes.batch_options(), precision=precision)
        b = make_batch(root, floor, opts, counts, pos)
        for _ in range(30):
            b.step()
        p = b.get_state("pos")
        for s in range(len(counts)):
            first, n = b.scene(s)
            o = O.splat_opts(opts["bakeResolution"], opts["bakeMipLevels"], opts["followSize"],
                             [float(v) for v in b.origin(s)], opts["bakeBand"], opts["surfaceRadius"])
            grids = O.splat_bake(o, p[first:first + n])
            for m in range(b.mips):
                got = b.codes(s, m)
                assert np.array_equal(got, grids[m].reshape(-1)), \
                    f"{precision} scene {s} mip {m}: {(got != grids[m].reshape(-1)).sum()} codes differ"
        b.destroy()


def test_tuning_does_not_change_results(root, floor):
    counts, pos = small_batch(seed=11)
    opts = scenes.batch_options()
    a = make_batch(root, floor, opts, counts, pos)
    t = make_batch(root, floor, opts, counts, pos)
    want = {"block": 128, "stage": 64, "chunk": 512, "groups_per_cu": 3, "fluid_xcd": 1, "splat_xcd": 1,
            "splat_owner": 1, "exact_lds": 0, "early_bin": 1}
    t.tuning = want
    assert t.tuning == want
    for _ in range(2):
        a.step()
        t.step()
    assert bits_equal(a.get_state("pos"), t.get_state("pos"))
    for s in range(len(counts)):
        for m in range(a.mips):
            assert np.array_equal(a.codes(s, m), t.codes(s, m))
    with pytest.raises(clayhip.ClayError):
        t.tuning = {"block": 96}
    a.destroy()
    t.destroy()


def test_fp16lds_batch_within_tolerance_of_single(root, floor):
    counts, pos = small_batch(seed=13)
    opts = dict(scenes.batch_options(), precision="fp16lds")
    b = make_batch(root, floor, opts, counts, pos)
    b.step()
    bp = b.get_state("pos")
    h = b.constants.h
    for s in (0, 3, 5):
        first, n = b.scene(s)
        sim = clayhip.FluidSim(root, floor, dict(opts, capacity=n))
        sim.set_state(pos[first:first + n])
        sim.step()
        d = np.abs(bp[first:first + n].astype(np.float64) - sim.get_state("pos")) / h
        assert np.sqrt(np.mean(d * d)) < 2e-4 and d.max() < 5e-2, (s, d.max())
        sim.destroy()
    b.destroy()


def test_batch_conserved_per_scene(root, floor):
    counts, pos = small_batch(seed=17)
    b = make_batch(root, floor, scenes.batch_options(), counts, pos)
    b.step()
    p, v = b.get_state("pos").astype(np.float64), b.get_state("vel").astype(np.float64)
    m = b.constants.mass
    for s in range(len(counts)):
        first, n = b.scene(s)
        c = b.conserved(s)
        assert c["mass"] == pytest.approx(m * n, rel=1e-15)
        assert np.allclose(c["momentum"], m * v[first:first + n].sum(axis=0), rtol=1e-12, atol=1e-12)
        assert np.allclose(c["com"], p[first:first + n].mean(axis=0), rtol=1e-12, atol=1e-12)
    b.destroy()
