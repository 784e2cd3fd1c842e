"""A FluidSim replaying its frames from a captured graph equals one stepped eagerly.

The 16k dam break is stepped twice side by side, eagerly and with clay_fluid_set_graph:
positions, velocities and every bake level must be bit-identical after each frame,
through a spawn (an eager frame that changes the live count, then a recapture), a
configure (the parameters are launch arguments of the captured frame) and a state
upload (the first sort and densities run eagerly again).
"""
import numpy as np
import pytest

import clayhip
from clayhip import scenes
from gpu_helpers import bits_equal, make_world

pytestmark = pytest.mark.gpu


def test_graph_replay_matches_eager_frames(root):
    sc = scenes.dam_break_small()
    opts = dict(sc["options"], capacity=sc["pos"].shape[0] + 4096)
    world, _, vol = make_world(root, sc)
    eager = clayhip.FluidSim(root, world, opts)
    graph = clayhip.FluidSim(root, world, opts)
    graph.set_graph(True)
    for s in (eager, graph):
        s.set_state(sc["pos"])

    def same(tag):
        for what in ("pos", "vel", "density"):
            a, b = eager.get_state(what), graph.get_state(what)
            assert bits_equal(a, b), f"{tag}: {what} differs"
        for m in range(opts["bakeMipLevels"]):
            assert np.array_equal(eager.surface.codes(m), graph.surface.codes(m)), f"{tag}: codes of mip {m} differ"

    for f in range(3):
        eager.step()
        graph.step()
        same(f"frame {f}")
    for s in (eager, graph):
        s.spawn(1000, (0.0, 3.5, 0.0), (0.0, -1.0, 0.0))
    for f in range(2):
        eager.step()
        graph.step()
        same(f"after spawn, frame {f}")
    assert eager.particleCount == graph.particleCount == sc["pos"].shape[0] + 1000
    for s in (eager, graph):
        s.configure({"maxSpeed": 9.0, "gravity": (0.5, -9.81, 0.0)})
    for f in range(2):
        eager.step()
        graph.step()
        same(f"after configure, frame {f}")
    for s in (eager, graph):
        s.set_state(sc["pos"])
    eager.step()
    graph.step()
    same("after set_state")
    eager.destroy()
    graph.destroy()
    world.destroy()
    if vol is not None:
        vol.destroy()
