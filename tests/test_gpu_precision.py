"""The FP16-LDS precision variant and the FP64 reductions, against the f64 oracle.

Tolerances, per substep from identical states: the plain path rounds like IEEE
f32, so its density agrees with f64 to 1e-5 (relative) and positions to 1e-4 h.
The FP16-LDS path stages neighbour offsets from their own cell centre as f16
(|o| <= h/2, rounding error <= 2^-16 ~ 1.5e-5 at h = 0.1), bounding the density
error at 2e-3 relative and the RMS position error at 2e-4 h.  Its maximum
position error is bounded loosely (0.05 h): the collision projection is
discontinuous at the contact threshold, and a particle whose velocity differs in
the last bits can land on the other side of it.
"""
import numpy as np
import pytest

import clayhip
from clayhip import scenes
from precision_report import run

pytestmark = pytest.mark.gpu


def test_per_substep_errors_against_f64(root):
    plain = run(root, "f32", steps=3)
    lds = run(root, "fp16lds", steps=3)
    p, q = plain["per_substep"], lds["per_substep"]
    assert p["density_rel"]["max"] < 1e-5
    assert p["position_over_h"]["max"] < 1e-4
    assert q["density_rel"]["max"] < 2e-3
    assert q["position_over_h"]["rms"] < 2e-4
    assert q["position_over_h"]["max"] < 5e-2
    assert q["density_rel"]["rms"] > p["density_rel"]["rms"]  # the staging is really f16


def test_conserved_quantities_in_fp64(root):
    sc = scenes.dam_break_small()
    world = clayhip.analyticField(root, sc["collider"]["brushes"])
    sim = clayhip.FluidSim(root, world, sc["options"])
    rng = np.random.default_rng(0)
    sim.set_state(sc["pos"], rng.normal(size=sc["pos"].shape).astype(np.float32))
    for _ in range(2):
        sim.step()
    c = sim.conserved()
    pos = sim.get_state("pos").astype(np.float64)
    vel = sim.get_state("vel").astype(np.float64)
    m = sim.constants.mass
    assert c["mass"] == pytest.approx(m * pos.shape[0], rel=1e-15)
    assert np.allclose(c["momentum"], m * vel.sum(axis=0), rtol=1e-12, atol=1e-12)
    assert np.allclose(c["com"], pos.mean(axis=0), rtol=1e-12, atol=1e-12)
    sim.destroy()
    world.destroy()
