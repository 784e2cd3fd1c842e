"""Particle-steps/sec including the SDF bake, on one MI355X.

A step is one FluidSim.step(): 2 substeps (bin + counting sort, density, forces +
integrate + collide) and the splat bake of every mip.  The default workload is
the 1M-particle dam break: 128 x 64 x 128 particles at spacing 0.05 over
[-3.2,3.2] x [0,3.2] x [-3.2,3.2], colliding with a 288^3 baked world volume,
baked every step into a 288^3 volume with 3 mips.  Inputs are resident in HBM
before the timed region starts.

Prints one JSON line (see the keys below).  --gpus N > 1 is refused: this build
measures one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "clay-engine_amd"))

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(kernel: str, particles: int, mips: int, live_voxels: int) -> float | None:
    """Compulsory HBM bytes of one launch of `kernel` (per-unit figures x units)."""
    per_particle = {
        "fluid_keys": 20,              # read pos 12, write key + id 8
        "fluid_reorder": 56,           # read key/id 8 + pos/vel 24, write sorted pos/vel 24
        "fluid_density": 16,           # read pos 12, write rho 4
        "fluid_forces_integrate": 52,  # read pos, vel, rho 28; write pos, vel 24
        "sort_hist": 4,
        "sort_scatter": 16,
    }
    if kernel in per_particle:
        return float(per_particle[kernel] * particles)
    if kernel == "splat_brick":          # all mips in one launch: position reads
        return float(12 * particles * mips)
    if kernel == "splat_plain":          # one launch per mip
        return float(12 * particles)
    if kernel == "splat_brick_resolve":  # 16 B per live voxel (code clear/resolve + texel)
        return float(16 * live_voxels)
    return None


def committed_traffic(kernel: str, workload: str) -> float | None:
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this workload (profiles/counters_<workload>.md), already
    scaled by the copy-kernel calibration; None when there is none."""
    path = ROOT / "profiles" / f"counters_{workload.replace('-', '_')}.md"
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        parts = line.split()
        if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == kernel:
            return float(parts[2])
    return None


def build_workload(root, name: str, splat: str, precision: str = "f32"):
    import clayhip
    from clayhip import scenes

    scene = {"dam-break-1m": scenes.dam_break_large, "dam-break-16k": scenes.dam_break_small}[name]()
    spec = scene["collider"]
    keep = []
    if spec["kind"] == "analytic":
        world = clayhip.analyticField(root, spec["brushes"])
    else:
        vol = clayhip.SdfVolume(root, spec["volume"])
        vol.bakeBrushes(spec["brushes"])
        keep.append(vol)
        world = clayhip.volumeField(vol)
    opts = dict(scene["options"], splatMode=splat, precision=precision)
    sim = clayhip.FluidSim(root, world, opts)
    sim.set_state(scene["pos"])
    return scene, sim, world, keep


def cpu_baseline(scene, budget_s: float, threads: int) -> dict:
    """The oracle (plain C, f32 mode, OpenMP) on this host's cores: whole steps of
    the same workload until `budget_s` has elapsed (at least one)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import numpy as np

    import oracle as O
    from clayhip import make_brush

    O.set_threads(threads)
    opts = scene["options"]
    spec = scene["collider"]
    if spec["kind"] == "analytic":
        col = O.Collider().add_analytic([make_brush(b) for b in spec["brushes"]])
    else:
        v = spec["volume"]
        tex = O.bake_brushes(v["resolution"], v["mipLevels"], v["worldSize"], v["origin"], v["band"],
                             [make_brush(b) for b in spec["brushes"]])
        col = O.Collider().add_volume(tex[0][..., 0], v["resolution"], v["origin"], v["worldSize"], v["band"])
    fl = O.Fluid(opts, scene["pos"], collider=col)
    so = O.splat_opts(opts["bakeResolution"], opts["bakeMipLevels"], opts["worldSize"], opts["origin"],
                      opts["bakeBand"], opts["surfaceRadius"])
    steps = 0
    t0 = time.perf_counter()
    while True:
        for _ in range(opts["substeps"]):
            fl.substep()
        grids = O.splat_bake(so, fl.get("pos"))
        O.splat_resolve(so, grids)
        steps += 1
        if time.perf_counter() - t0 >= budget_s:
            break
    dt = time.perf_counter() - t0
    n = scene["pos"].shape[0]
    return {
        "value": n * steps / dt,
        "unit": "particle-steps/s",
        "cores": O.threads(),
        "kind": "port",
        "sample": f"{steps} full step(s) of the same workload ({n} particles, 2 substeps + "
                  f"{opts['bakeMipLevels']}-mip bake), {dt:.1f} s wall, oracle f32 mode with OpenMP",
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="dam-break-1m", choices=["dam-break-1m", "dam-break-16k"])
    ap.add_argument("--splat", default="tiled", choices=["tiled", "plain"])
    ap.add_argument("--precision", default="f32", choices=["f32", "fp16lds"],
                    help="fluid kernels: IEEE f32 (bit-exact) or FP16-LDS staged neighbours")
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of oracle work for cpu_baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 256 MiB device-to-device copy 5 times (HBM counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch

        src = torch.empty(64 * 1024 * 1024, dtype=torch.float32, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_launch": src.numel() * 4, "launches": 5}))
        return 0
    world_size = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world_size != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        return 2

    import numpy as np
    import torch

    import clayhip

    root = clayhip.Root(0)
    scene, sim, world, keep = build_workload(root, args.workload, args.splat, args.precision)
    n = scene["pos"].shape[0]
    mips = scene["options"]["bakeMipLevels"]
    stream = torch.cuda.Stream()
    for _ in range(args.warmup):
        sim.step(stream)
    stream.synchronize()
    root.synchronize()

    root.reset_stats()
    root.set_timing(True)
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    start.record(stream)
    for _ in range(args.steps):
        sim.step(stream)
    stop.record(stream)
    stream.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    root.set_timing(False)
    ms = start.elapsed_time(stop)
    ms_per_step = ms / args.steps
    stats = root.kernel_stats()

    live_voxels = sum(int(sim.surface.liveTiles(m).sum()) * 512 for m in range(mips))
    dom_name, (dom_ms, dom_n) = max(stats.items(), key=lambda kv: kv[1][0])
    dom_avg_s = dom_ms / dom_n / 1e3
    dom_bytes = algorithmic_bytes(dom_name, n, mips, live_voxels)
    roofline = None
    if dom_bytes is not None:
        achieved = dom_bytes / dom_avg_s / 1e9
        traffic = committed_traffic(dom_name, args.workload) if args.splat == "tiled" else None
        roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                    "kernel": dom_name, "avg_ms": round(dom_avg_s * 1e3, 4), "bytes_per_launch": dom_bytes}
    step_bytes = n * (2 * 124 + 12 * mips) + 16 * live_voxels
    line = {
        "metric": "particle-steps/sec incl. SDF bake (1 MI355X)",
        "value": n * args.steps / (ms / 1e3),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded jittered lattice)",
        "config": {
            "workload": scene["name"],
            "particles": n,
            "substeps": scene["options"]["substeps"],
            "bake": f"{scene['options']['bakeResolution']}^3 x {mips} mips",
            "splat": args.splat,
            "precision": args.precision,
            "live_voxels": live_voxels,
        },
        "step_roofline": {"bytes_per_step": step_bytes,
                          "achieved_GBs": round(step_bytes / (ms_per_step / 1e3) / 1e9, 2),
                          "frac": round(step_bytes / (ms_per_step / 1e3) / 1e9 / HBM_PEAK_GBS, 5)},
        "roofline": roofline,
        "kernels_ms_per_step": {k: round(v[0] / args.steps, 4) for k, v in sorted(stats.items(), key=lambda kv: -kv[1][0])},
        "wall_s": round(wall, 3),
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(scene, args.cpu_budget, min(16, os.cpu_count() or 1))
    else:
        line["cpu_baseline"] = None
    sim.destroy()
    world.destroy()
    for k in keep:
        k.destroy()
    root.destroy()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
