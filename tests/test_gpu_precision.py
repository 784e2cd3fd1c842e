"""The FP16-LDS precision variant and the FP64 reductions, against the f64 oracle.

Tolerances, per substep from identical states: the plain path rounds like IEEE
f32, so its density agrees with f64 to 1e-5 (relative) and positions to 1e-4 h.
The FP16-LDS path stages neighbour offsets from their own cell centre as f16
(|o| <= h/2, rounding error <= 2^-16 ~ 1.5e-5 at h = 0.1), bounding the density
error at 2e-3 relative and the RMS position error at 2e-4 h.  Its maximum
position error is bounded at 1e-2 h: the collision projection is discontinuous
at the contact threshold, and a particle whose velocity differs in the last bits
can land on the other side of it.  A pair closer than h/10 takes its offset from
the f32 positions, since f16 offsets would set its direction.
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
    assert q["position_over_h"]["max"] < 1e-2
    assert q["density_rel"]["rms"] > p["density_rel"]["rms"]  # the staging is really f16


def test_fp16lds_error_through_impact(root):
    """Frames 16-47 of the 16k dam break, where the column hits the floor and piles
    up (densities to more than twice the rest density): every substep, the f64
    oracle takes one substep from the GPU's own state.  FP16-LDS keeps the RMS
    position error below 1e-4 h, the largest below 5e-3 h (near-coincident pairs
    take their offsets in f32) and the density error below 1e-4 RMS; the collision
    branch taken from the GPU's and the oracle's velocities must agree for all but
    0.05% of the particles."""
    import oracle as O

    sc = scenes.dam_break_small()
    brushes = sc["collider"]["brushes"]
    col = O.Collider().add_analytic([clayhip.make_brush(b) for b in brushes])
    world = clayhip.analyticField(root, brushes)
    worst = {}
    for precision in ("f32", "fp16lds"):
        opts = dict(sc["options"], precision=precision)
        sim = clayhip.FluidSim(root, world, opts)
        sim.set_state(sc["pos"])
        h, rad, dt = sim.constants.h, sim.constants.radius, sim.constants.dt
        for _ in range(16):
            sim.step()
        sq, sq_rho, count, emax, flips, rho_top = 0.0, 0.0, 0, 0.0, 0, 0.0
        for frame in range(16, 48):
            for _ in range(opts["substeps"]):
                p0, v0 = sim.get_state("pos"), sim.get_state("vel")
                ref = O.Fluid(opts, p0.astype(np.float64), v0.astype(np.float64), collider=col, precision="f64")
                ref.substep()
                sim.substep()
                e = np.linalg.norm(sim.get_state("pos") - ref.get("pos"), axis=1) / h
                rho_ref = ref.get("density")
                er = (sim.get_state("density") - rho_ref) / rho_ref
                dw_g, _ = col.eval(p0 + sim.get_state("nvel").astype(np.float64) * dt, precision="f64")
                dw_o, _ = col.eval(p0 + ref.get("nvel") * dt, precision="f64")
                f = int(((dw_g < rad) != (dw_o < rad)).sum())
                assert f <= 5e-4 * e.size, f"{precision} frame {frame}: {f} collision flips"
                flips += f
                sq += float((e * e).mean())
                sq_rho += float((er * er).mean())
                count += 1
                emax = max(emax, float(e.max()))
                rho_top = max(rho_top, float(rho_ref.max()))
        worst[precision] = (np.sqrt(sq / count), emax, np.sqrt(sq_rho / count), flips)
        assert rho_top > 2.0 * sim.constants.rest_density, "the column must pile up"
        sim.destroy()
    world.destroy()
    print({p: dict(zip(("position_over_h_rms", "position_over_h_max", "density_rel_rms", "collision_flips"), w))
           for p, w in worst.items()})
    rms, emax, rho_rms, _ = worst["f32"]
    assert rms < 2e-6 and emax < 2e-5 and rho_rms < 1e-6, worst
    rms, emax, rho_rms, _ = worst["fp16lds"]
    assert rms < 1e-4 and emax < 5e-3 and rho_rms < 1e-4, worst


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


def test_near_coincident_pair_takes_its_direction_from_f32(root):
    """Two particles 1.3e-3 h apart inside a block at rest, no gravity: their pair
    pushes them apart along their offset.  Staged as f16 offsets from the cell
    centre (rounding up to ~1.5e-4 h per component) that offset would be off by
    tens of percent; the forces pass takes pairs closer than h/10 from the f32
    positions, so both particles' staged velocities match the f64 oracle's to 1%."""
    import oracle as O

    pos = scenes.lattice(10, 10, 10, 0.05, (-0.25, 0.5, -0.25), seed=3)
    k = int(np.argmin(np.linalg.norm(pos - pos.mean(axis=0), axis=1)))
    extra = pos[k] + np.array([1.0e-4, 0.7e-4, 0.3e-4], dtype=np.float32)
    pos = np.vstack([pos, extra[None, :]]).astype(np.float32)
    brushes = [scenes.floor_brush()]
    opts = dict(scenes.dam_break_options(pos.shape[0], 32, 1), precision="fp16lds", gravity=(0.0, 0.0, 0.0))
    world = clayhip.analyticField(root, brushes)
    sim = clayhip.FluidSim(root, world, opts)
    sim.set_state(pos)
    sim.substep()
    got = sim.get_state("nvel").astype(np.float64)
    col = O.Collider().add_analytic([clayhip.make_brush(b) for b in brushes])
    ref = O.Fluid(opts, pos.astype(np.float64), collider=col, precision="f64")
    ref.substep()
    want = ref.get("nvel")
    sim.destroy()
    world.destroy()
    for i in (k, pos.shape[0] - 1):
        assert np.linalg.norm(want[i]) > 0.1, "the pair must repel"
        err = np.linalg.norm(got[i] - want[i]) / np.linalg.norm(want[i])
        assert err < 1e-2, f"particle {i}: relative nvel error {err:.3g}"
