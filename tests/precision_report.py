"""Error of the GPU fluid paths against the f64 oracle on the 16k dam break.

For every substep of a 100-step run, the f64 oracle starts from the GPU's own
state (identical inputs) and takes one substep; density, force (the staged
velocity `nvel`) and position are compared.  Separately, the free-running GPU
trajectory is compared with a free-running f64 trajectory after 1, 2, 5, ... N
steps.  Errors report max, RMS and the 99.9th percentile: the collision
projection is discontinuous at the contact threshold, so a few particles carry
errors far above the bulk.

Usage (on a GPU): python tests/precision_report.py [--steps 100]
Prints a JSON object; tests/test_gpu_precision.py runs a short version.
"""
from __future__ import annotations

import argparse
import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT / "clay-engine_amd"))
sys.path.insert(0, str(ROOT / "oracle"))

import clayhip  # noqa: E402
import oracle as O  # noqa: E402
from clayhip import scenes  # noqa: E402


CHECKPOINTS = (1, 2, 5, 10, 20, 50, 100)


def _stats(err: np.ndarray) -> tuple[float, float, float]:
    """max |e|, RMS and the 99.9th percentile of |e|."""
    err = np.abs(np.asarray(err, dtype=np.float64))
    return float(np.max(err)), float(np.sqrt(np.mean(err * err))), float(np.quantile(err, 0.999))


def run(root, precision: str, steps: int, free_steps: int | None = None) -> dict:
    sc = scenes.dam_break_small()
    opts = dict(sc["options"], precision=precision)
    brushes = sc["collider"]["brushes"]
    recs = [clayhip.make_brush(b) for b in brushes]
    col = O.Collider().add_analytic(recs)
    world = clayhip.analyticField(root, brushes)
    sim = clayhip.FluidSim(root, world, opts)
    sim.set_state(sc["pos"])
    h = sim.constants.h
    worst = {"density_rel": [0.0, 0.0], "force": [0.0, 0.0], "position_over_h": [0.0, 0.0]}  # max, p99.9
    sums = {k: 0.0 for k in worst}
    count = 0
    substeps = opts["substeps"]
    cons0 = sim.conserved()
    for _ in range(steps):
        for _ in range(substeps):
            pos0, vel0 = sim.get_state("pos"), sim.get_state("vel")
            ref = O.Fluid(opts, pos0.astype(np.float64), vel0.astype(np.float64), collider=col, precision="f64")
            ref.substep()
            sim.substep()
            rho_ref = ref.get("density")
            e_rho = (sim.get_state("density") - rho_ref) / rho_ref
            e_f = (sim.get_state("nvel") - ref.get("nvel")).reshape(-1)
            e_p = (sim.get_state("pos") - ref.get("pos")).reshape(-1) / h
            for key, e in (("density_rel", e_rho), ("force", e_f), ("position_over_h", e_p)):
                mx, rms, p999 = _stats(e)
                worst[key][0] = max(worst[key][0], mx)
                worst[key][1] = max(worst[key][1], p999)
                sums[key] += rms * rms
            count += 1
    out = {"precision": precision, "substeps_compared": count,
           "per_substep": {k: {"max": worst[k][0], "p99.9": worst[k][1],
                               "rms": float(np.sqrt(sums[k] / max(count, 1)))} for k in worst}}
    cons1 = sim.conserved()
    out["conserved_fp64"] = {"mass_before": cons0["mass"], "mass_after": cons1["mass"],
                             "momentum_after": cons1["momentum"], "com_after": cons1["com"]}
    sim.destroy()
    if free_steps:
        sim = clayhip.FluidSim(root, world, opts)
        sim.set_state(sc["pos"])
        ref = O.Fluid(opts, sc["pos"].astype(np.float64), collider=col, precision="f64")
        curve = []
        for step in range(1, free_steps + 1):
            sim.step()
            for _ in range(substeps):
                ref.substep()
            if step in CHECKPOINTS or step == free_steps:
                d = (sim.get_state("pos") - ref.get("pos")) / h
                mx, rms, _ = _stats(np.linalg.norm(d, axis=1))
                curve.append({"step": step, "position_over_h_max": mx, "position_over_h_rms": rms})
        out["free_run"] = curve
        sim.destroy()
    world.destroy()
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=100)
    args = ap.parse_args()
    O.set_threads(16)
    root = clayhip.Root(0)
    res = [run(root, p, args.steps, free_steps=args.steps) for p in ("f32", "fp16lds")]
    root.destroy()
    print(json.dumps(res, indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())
