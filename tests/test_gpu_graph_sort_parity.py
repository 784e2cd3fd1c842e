"""Captured frames stay bit-identical to eager frames whatever runs between replays.

A frame sorts its particles several times (one stable radix sort per substep plus
the frame-end sort), each sort running three digit passes over two buffer pairs.  A
captured frame holds the buffer addresses it was captured with, so the sort must
leave its result in the same buffers however many passes and sorts ran before a
replay.  These cases run an odd number of sorts between replays: three substeps per
frame (four sorts), a state upload of the same particle count after replays (the
graph is kept and one eager sort runs before the next replay), and an extra eager
substep between frames.
"""
import numpy as np
import pytest

import clayhip
from clayhip import scenes
from gpu_helpers import bits_equal, describe_diff

pytestmark = pytest.mark.gpu


def _same(tag, a, b, mips, codes_a, codes_b):
    for what in ("pos", "vel", "density"):
        x, y = a.get_state(what), b.get_state(what)
        assert bits_equal(x, y), describe_diff(f"{tag}: {what}", x, y)
    for m in range(mips):
        assert np.array_equal(codes_a(m), codes_b(m)), f"{tag}: codes of mip {m} differ"


def test_fluidsim_graph_with_three_substeps_and_same_count_reset(root):
    sc = scenes.dam_break_small()
    opts = dict(sc["options"], substeps=3)
    world = clayhip.analyticField(root, [scenes.floor_brush()])
    eager = clayhip.FluidSim(root, world, opts)
    graph = clayhip.FluidSim(root, world, opts)
    graph.set_graph(True)
    for s in (eager, graph):
        s.set_state(sc["pos"])
    mips = opts["bakeMipLevels"]

    def check(tag):
        _same(tag, eager, graph, mips, lambda m: eager.surface.codes(m), lambda m: graph.surface.codes(m))

    for f in range(3):
        eager.step()
        graph.step()
        check(f"frame {f}")
    # Same live count: the graph is kept, one eager sort and density run first.
    moved = sc["pos"] + np.float32(0.01)
    for s in (eager, graph):
        s.set_state(moved)
    for f in range(2):
        eager.step()
        graph.step()
        check(f"after set_state of the same count, frame {f}")
    for s in (eager, graph):
        s.substep()
    for f in range(2):
        eager.step()
        graph.step()
        check(f"after an eager substep, frame {f}")
    eager.destroy()
    graph.destroy()
    world.destroy()


def test_batch_graph_with_three_substeps_and_reset(root):
    counts = np.array([1500, 3001, 800, 5000], dtype=np.uint32)
    blocks = [scenes.drop_block(int(c), d, seed=700 + i) for i, (c, d) in enumerate(zip(counts, [0.5, 2.0, 1.2, 0.7]))]
    pos = np.concatenate(blocks, axis=0)
    world = clayhip.analyticField(root, [scenes.floor_brush()])
    for precision in ("f32", "fp16lds"):
        opts = dict(scenes.batch_options(), substeps=3, precision=precision)
        g = clayhip.FluidBatch(root, world, opts, counts)
        e = clayhip.FluidBatch(root, world, opts, counts)
        g.set_graph(True)
        e.set_graph(False)
        for b in (g, e):
            b.set_state(pos)

        def check(tag):
            for what in ("pos", "vel", "density"):
                x, y = g.get_state(what), e.get_state(what)
                assert bits_equal(x, y), describe_diff(f"{precision} {tag}: {what}", x, y)
            for s in range(len(counts)):
                for m in range(g.mips):
                    assert np.array_equal(g.codes(s, m), e.codes(s, m)), f"{precision} {tag}: scene {s} mip {m}"

        for f in range(3):
            g.step()
            e.step()
            check(f"frame {f}")
        for b in (g, e):
            b.set_state(pos)
        for f in range(3):
            g.step()
            e.step()
            check(f"after a reset, frame {f}")
        g.destroy()
        e.destroy()
    world.destroy()
