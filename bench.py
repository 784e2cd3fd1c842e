"""Particle-steps/sec including the SDF bake, on one MI355X.

A step is one FluidSim.step(): 2 substeps (bin + sort, density, forces +
integrate + collide) and the splat bake of every mip.  Inputs are resident in
HBM before the timed region starts.

Default workload (`batched-256`): 256 independent dam-break scenes with particle
counts log-uniform in [1024, 65536] (3 532 015 in total) and drop heights in
[0.5, 3.0], each in a 6.4-unit world over the roundBox floor, each baking a
32^3 x 3-mip volume 3.2 units wide that follows its particles.  The measured
path is the batch with FP16-LDS staging, one captured graph per frame and the
autotuned launch shapes.  Beside it, in the same run:
  * the same scenes stepped one FluidSim at a time (plain f32 path),
  * the single-scene 1M dam break (`dam-break-1m`, plain f32 path),
  * the position error of the FP16-LDS batch against the f32 batch over one step
    from the same state,
  * the CPU oracle on a bounded sample of the scenes.
`--workload dam-break-1m` / `dam-break-16k` measure a single FluidSim instead,
`--workload coupled` the coupled frame (fluid + ClaySolver body + eroding world per
scene), and `--workload wire` the baked volume's stream format (pack + unpack of the
288^3 x 3 world; its own metric, texels/s).

Prints one JSON line.  --gpus N > 1 is refused: this build measures one GPU.
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
METRIC = "particle-steps/sec incl. SDF bake (1 MI355X); batched+FP16-LDS vs plain at stated err"


def algorithmic_bytes(kernel: str, particles: int, mips: int, live_voxels: int) -> float | None:
    """Compulsory HBM bytes of one launch of `kernel` (per-unit figures x units)."""
    kernel = kernel.removesuffix("_lds")
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


# VALU issue: MI355X's published FP32 vector peak, 157.3 TFLOP/s, is two wave64
# FMA instructions (128 flops) per CU per cycle at 2.4 GHz.  gfx950 reaches that
# issue rate only for two-operand VGPR forms (v_fmac_f32, v_add_f32, v_mul_f32,
# bitwise, moves); three-operand, compare/select and max forms issue at about half
# and transcendentals at a quarter of it.  clay_valu_peak measures each form's rate
# on the device this run is on; the fastest is reported beside the published figure.
VALU_PEAK_PUBLISHED = 157.3e12 / 128  # wave64 VALU instructions per second
_VALU_MEASURED: dict = {}


def valu_measured(root) -> dict:
    """Wave64 instructions per second this device retires for each instruction form
    clay_valu_peak knows ("fmac" is the full-rate ceiling), measured once per run."""
    if not _VALU_MEASURED:
        for f in root.VALU_FORMS:
            _VALU_MEASURED[f] = max(root.valu_peak(f) for _ in range(2))
    return _VALU_MEASURED


def committed_valu(kernel: str, workload: str) -> float | None:
    """SQ_INSTS_VALU per launch of `kernel` (wave instructions) from the committed
    rocprofv3 counter summary of this workload, or None."""
    path = ROOT / "profiles" / f"counters_{workload.replace('-', '_')}.md"
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        parts = line.split()
        if len(parts) == 3 and parts[0] == "valu_insts_per_launch" and parts[1] == kernel:
            return float(parts[2])
    return None


def roofline_of(root, stats: dict, particles: int, mips: int, live_voxels: int, workload: str) -> dict | None:
    """Roofline object of the dominant kernel of a timed pass: its algorithmic HBM
    bytes per launch over its average launch time against the HBM peak, and beside
    it the fraction of the VALU issue peak its counted wave instructions
    (SQ_INSTS_VALU per launch) take over the same average time, against both the
    published FP32 issue rate and the full-rate ceiling measured in this run."""
    if not stats:
        return None
    name, (tot_ms, launches) = max(stats.items(), key=lambda kv: kv[1][0])
    avg_s = tot_ms / launches / 1e3
    nbytes = algorithmic_bytes(name, particles, mips, live_voxels)
    if nbytes is None:
        return None
    achieved = nbytes / avg_s / 1e9
    out = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": committed_traffic(name, workload),
           "kernel": name, "avg_ms": round(avg_s * 1e3, 4), "bytes_per_launch": nbytes}
    valu = valu_object(root, committed_valu(name, workload), avg_s)
    if valu is not None:
        out["valu"] = valu
    return out


def valu_object(root, valu: float | None, avg_s: float) -> dict | None:
    """The VALU side of a roofline: `valu` wave instructions per launch over `avg_s`
    against the published issue rate and the fastest form measured in this run."""
    if valu is None:
        return None
    forms = valu_measured(root)
    return {"insts_per_launch": valu, "peak_insts_per_s": VALU_PEAK_PUBLISHED,
            "frac": round(valu / (VALU_PEAK_PUBLISHED * avg_s), 4),
            "measured_peak_insts_per_s": max(forms.values()),
            "frac_of_measured": round(valu / (max(forms.values()) * avg_s), 4),
            "measured_insts_per_s_by_form": {f: float(f"{r:.4g}") for f, r in forms.items()}}


def step_roofline(particles: int, substeps: int, mips: int, live_voxels: int, ms_per_step: float) -> dict:
    b = particles * (substeps * 124 + 12 * mips) + 16 * live_voxels
    gbs = b / (ms_per_step / 1e3) / 1e9
    return {"bytes_per_step": b, "achieved_GBs": round(gbs, 2), "frac": round(gbs / HBM_PEAK_GBS, 5)}


def timed(step, steps: int, warmup: int, stream) -> float:
    """ms per step of `step(stream)` over `steps` steps after `warmup`, device-bracketed."""
    import torch

    for _ in range(warmup):
        step(stream)
    stream.synchronize()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(steps):
        step(stream)
    stop.record(stream)
    stream.synchronize()
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / steps


def kernel_stats(root, step, steps: int, stream) -> dict:
    """Per-kernel HIP-event totals over `steps` eager steps: {name: (ms, launches)}."""
    root.synchronize()
    root.reset_stats()
    root.set_timing(True)
    for _ in range(steps):
        step(stream)
    stream.synchronize()
    root.synchronize()
    root.set_timing(False)
    return root.kernel_stats()


def per_step(stats: dict, steps: int) -> dict:
    return {k: round(v[0] / steps, 4) for k, v in sorted(stats.items(), key=lambda kv: -kv[1][0])}


# ---- single FluidSim workloads ------------------------------------------------
def build_single(root, name: str, splat: str, precision: str = "f32"):
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


def measure_single(root, name: str, splat: str, precision: str, steps: int, warmup: int, stream,
                   autotune_now: bool = False) -> dict:
    from clayhip import autotune

    scene, sim, world, keep = build_single(root, name, splat, precision)
    n = scene["pos"].shape[0]
    tune_src = "default"
    if splat == "tiled":
        key = autotune.cache_key(name, precision)
        if autotune_now:
            autotune.autotune(sim, key, frames=3, space=autotune.space_for(precision),
                              state=lambda: sim.set_state(scene["pos"]))
            tune_src = "autotuned in this run"
        else:
            cached = autotune.load(autotune.DEFAULT_CACHE, key)
            if cached:
                sim.tuning = cached
                tune_src = "autotune cache"
        sim.set_state(scene["pos"])
    mips = scene["options"]["bakeMipLevels"]
    ms = timed(sim.step, steps, warmup, stream)
    ks = min(steps, 5)
    stats = kernel_stats(root, sim.step, ks, stream)
    atomics = splat_atomics(root, sim.splat_atomics(), n, stats) if splat == "tiled" else None
    live = sum(int(sim.surface.liveTiles(m).sum()) * 512 for m in range(mips))
    out = {
        "workload": scene["name"], "particles": n, "ms_per_step": ms, "value": n / (ms / 1e3),
        "substeps": scene["options"]["substeps"], "mips": mips, "live_voxels": live,
        "kernels_ms_per_step": per_step(stats, ks),
        "roofline": roofline_of(root, stats, n, mips, live, name) if splat == "tiled" else None,
        "splat_atomics": atomics,
        "tuning": dict(sim.tuning, source=tune_src),
        "scene": scene,
    }
    sim.destroy()
    world.destroy()
    for k in keep:
        k.destroy()
    return out


def cpu_protocol(step, budget_s: float) -> dict:
    """The CPU timing protocol (5 warm-up steps, then 3 runs of 100 timed steps,
    median run) within a time budget: warm-up steps (5, or
    fewer when one step takes more than 3% of the budget), then 3 runs of K timed
    steps (100, or as many as a third of the budget holds), wall clock around each
    run; the median run's seconds per step."""
    import statistics

    t0 = time.perf_counter()
    step()
    first = time.perf_counter() - t0
    warm = min(5, max(1, int(0.15 * budget_s / max(first, 1e-9))))
    for _ in range(warm - 1):
        step()
    k = max(1, min(100, int(budget_s / (3 * max(first, 1e-9)))))
    runs = []
    for _ in range(3):
        t0 = time.perf_counter()
        for _ in range(k):
            step()
        runs.append((time.perf_counter() - t0) / k)
    return {"s_per_step": statistics.median(runs), "warmup": warm, "timed": k, "runs": [round(r, 4) for r in runs]}


def cpu_baseline_single(scene, budget_s: float, threads: int) -> dict:
    """The oracle (plain C, f32 mode, OpenMP) on this host's cores: whole steps of
    the same workload until `budget_s` has elapsed (at least one)."""
    sys.path.insert(0, str(ROOT / "oracle"))
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

    def step():
        for _ in range(opts["substeps"]):
            fl.substep()
        O.splat_resolve(so, O.splat_bake(so, fl.get("pos")))

    pr = cpu_protocol(step, budget_s)
    n = scene["pos"].shape[0]
    return {
        "value": n / pr["s_per_step"], "unit": "particle-steps/s", "cores": O.threads(), "kind": "port",
        "sample": f"the whole workload ({n} particles, 2 substeps + {opts['bakeMipLevels']}-mip bake per step): "
                  f"{pr['warmup']} warm-up step(s), then 3 runs of {pr['timed']} timed step(s) (s/step "
                  f"{pr['runs']}), median run; oracle f32 mode with OpenMP. Fewer than the protocol's 5 + 100 "
                  f"steps where the {budget_s:.0f} s budget holds fewer",
    }


# ---- batched workload ----------------------------------------------------------
def cpu_baseline_batched(scene, budget_s: float, threads: int) -> dict:
    """The oracle on whole scenes of the batch, one after another, each taking one
    step (2 substeps, then the splat of its following volume centred as the device
    centres it), until `budget_s` has elapsed."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import numpy as np

    import oracle as O
    from clayhip import make_brush

    O.set_threads(threads)
    opts = scene["options"]
    col = O.Collider().add_analytic([make_brush(b) for b in scene["collider"]["brushes"]])
    counts = scene["counts"]
    base = np.concatenate([np.zeros(1, np.int64), np.cumsum(counts, dtype=np.int64)])
    half = opts["followSize"] / 2
    picked = list(range(0, len(counts), max(1, len(counts) // 8)))[:8]  # spread over the count range
    fluids = [O.Fluid(dict(opts, capacity=int(counts[s])), scene["pos"][base[s]:base[s + 1]], collider=col)
              for s in picked]

    def step():
        for fl in fluids:
            for _ in range(opts["substeps"]):
                fl.substep()
            p = fl.get("pos")
            centre = (p.min(axis=0) + p.max(axis=0)) * np.float32(0.5)
            origin = [float(np.float32(c - half)) for c in centre.astype(np.float64)]
            so = O.splat_opts(opts["bakeResolution"], opts["bakeMipLevels"], opts["followSize"], origin,
                              opts["bakeBand"], opts["surfaceRadius"])
            O.splat_resolve(so, O.splat_bake(so, p))

    pr = cpu_protocol(step, budget_s)
    done = int(sum(int(counts[s]) for s in picked))
    return {
        "value": done / pr["s_per_step"], "unit": "particle-steps/s", "cores": O.threads(), "kind": "port",
        "sample": f"scenes {picked} of the batch ({done} particles; a step is 2 substeps + the "
                  f"{opts['bakeMipLevels']}-mip following bake of each): {pr['warmup']} warm-up step(s), then 3 "
                  f"runs of {pr['timed']} timed step(s) (s/step {pr['runs']}), median run; oracle f32 mode with "
                  f"OpenMP. Fewer than the protocol's 5 + 100 steps where the {budget_s:.0f} s budget holds fewer",
    }


def measure_per_scene(root, world, scene, steps: int, stream) -> dict:
    """The same scenes, one f32 FluidSim each, stepped one after another."""
    import numpy as np

    import clayhip

    counts = scene["counts"]
    base = np.concatenate([np.zeros(1, np.int64), np.cumsum(counts, dtype=np.int64)])
    sims = []
    for s in range(len(counts)):
        sim = clayhip.FluidSim(root, world, dict(scene["options"], capacity=int(counts[s])))
        sim.set_state(scene["pos"][base[s]:base[s + 1]])
        sims.append(sim)

    def frame(st):
        for sim in sims:
            sim.step(st)

    ms = timed(frame, steps, 1, stream)
    for sim in sims:
        sim.destroy()
    n = int(counts.sum())
    return {"path": "one f32 FluidSim per scene, same scenes", "ms_per_step": ms, "value": n / (ms / 1e3),
            "steps": steps}


def splat_atomics(root, counts: tuple[int, int], particles: int, stats: dict) -> dict:
    """The owner walk's atomics per bake (counted by a tallying re-bake of the same
    state) against the measured issue rate of ds_min_u32, the splat's roof."""
    lds, glob = counts
    tot_ms, launches = stats.get("splat_brick", (0.0, 0))
    avg_s = tot_ms / launches / 1e3 if launches else float("nan")
    peak_lds = root.atomic_peak("lds_min")
    peak_glob = root.atomic_peak("global_min")
    achieved = lds / avg_s
    return {"lds_min_per_bake": lds, "global_min_per_bake": glob, "lds_min_per_particle": lds / particles,
            "splat_brick_avg_ms": round(avg_s * 1e3, 4), "achieved_lds_min_per_s": achieved,
            "peak_lds_min_per_s": peak_lds, "frac": round(achieved / peak_lds, 4),
            "peak_global_min_per_s": peak_glob,
            "note": "peaks measured in this run by clay_atomic_peak (conflict-free ds_min_u32; global atomicMin, "
                    "a wave per 256-byte span)"}


def measure_batched_1m(root, args, stream, tuning: dict) -> dict:
    """The batched path at 1M particles in total: 64 scenes of 16 384 particles
    (1 048 576 in total), batched with FP16-LDS staging and a captured graph, with
    the batched workload's tuned launch shapes."""
    import clayhip
    from clayhip import scenes

    scene = scenes.batched_even(64, 16384)
    world = clayhip.analyticField(root, scene["collider"]["brushes"])
    b = clayhip.FluidBatch(root, world, dict(scene["options"], precision=args.precision), scene["counts"])
    b.set_graph(not args.no_graph)
    b.tuning = tuning
    b.set_state(scene["pos"])
    ms = timed(b.step, args.steps, args.warmup, stream)
    n = int(scene["counts"].sum())
    b.destroy()
    world.destroy()
    return {"path": f"batched, {args.precision}, graph, " + scene["name"], "particles": n, "ms_per_step": ms,
            "value": n / (ms / 1e3)}


def measure_batched(root, args, stream) -> dict:
    import numpy as np

    import clayhip
    from clayhip import autotune, scenes

    scene = scenes.batched(256)
    world = clayhip.analyticField(root, scene["collider"]["brushes"])
    opts = dict(scene["options"], precision=args.precision)
    b = clayhip.FluidBatch(root, world, opts, scene["counts"])
    b.set_graph(not args.no_graph)
    b.set_state(scene["pos"])
    key = autotune.cache_key("batched-256", args.precision)
    tune_src = "default"
    if args.autotune:
        autotune.autotune(b, key, frames=3, space=autotune.space_for(args.precision),
                          state=lambda: b.set_state(scene["pos"]))
        tune_src = "autotuned in this run"
    else:
        cached = autotune.load(autotune.DEFAULT_CACHE, key)
        if cached:
            b.tuning = cached
            tune_src = "autotune cache"
    b.set_state(scene["pos"])
    n = int(scene["counts"].sum())
    mips = opts["bakeMipLevels"]
    ms = timed(b.step, args.steps, args.warmup, stream)
    # Per-kernel breakdown from eager frames (a graph replay has no per-launch events).
    b.set_graph(False)
    stats = kernel_stats(root, b.step, 3, stream)
    atomics = splat_atomics(root, b.splat_atomics(), n, stats)
    live = sum(int(b.live_tiles(s, m).sum()) * 512 for s in range(b.scenes) for m in range(mips))
    # Error of the measured path against the f32 batch over one step from its current state.
    err = None
    if args.precision != "f32":
        pos, vel = b.get_state("pos"), b.get_state("vel")
        ref = clayhip.FluidBatch(root, world, dict(opts, precision="f32"), scene["counts"])
        ref.tuning = b.tuning
        ref.set_graph(False)
        for x in (b, ref):
            x.set_state(pos, vel)
            x.step(stream)
        stream.synchronize()
        d = np.linalg.norm(b.get_state("pos").astype(np.float64) - ref.get_state("pos"), axis=1) / b.constants.h
        err = {"vs": "f32 batch, one step from the same state", "position_over_h_rms": float(np.sqrt(np.mean(d * d))),
               "position_over_h_p999": float(np.quantile(d, 0.999)), "position_over_h_max": float(d.max())}
        ref.destroy()
    out = {
        "workload": scene["name"], "particles": n, "scenes": b.scenes, "ms_per_step": ms,
        "value": n / (ms / 1e3), "substeps": opts["substeps"], "mips": mips, "live_voxels": live,
        "tuning": dict(b.tuning, source=tune_src), "graph": not args.no_graph,
        "kernels_ms_per_step": per_step(stats, 3),
        "roofline": roofline_of(root, stats, n, mips, live, "batched-256"),
        "splat_atomics": atomics,
        "error": err, "scene": scene,
    }
    tuning = {k: v for k, v in b.tuning.items()}
    b.destroy()
    if not args.no_compare:
        out["plain_per_scene"] = measure_per_scene(root, world, scene, min(args.steps, 5), stream)
        out["batched_1m"] = measure_batched_1m(root, args, stream, tuning)
    world.destroy()
    return out


def measure_coupled(root, args, stream) -> dict:
    """The coupled frame (fluid + ClaySolver body + eroding world per scene), timed
    with FP16-LDS staging and with the plain f32 path on the same scenes."""
    import numpy as np

    from clayhip.coupled import CoupledBatch, coupled_scene

    sc = coupled_scene(scenes=args.scenes)
    out = {"workload": sc["name"], "particles": int(sc["counts"].sum()), "scenes": args.scenes,
           "substeps": sc["options"]["substeps"], "mips": sc["options"]["bakeMipLevels"], "scene": sc}
    ends = {}
    for precision in (args.precision, "f32") if args.precision != "f32" else ("f32",):
        cb = CoupledBatch(root, sc["counts"], sc["pos"], dict(sc["options"], precision=precision), sc["world"],
                          sc["world_brushes"])
        ms = timed(cb.step, args.steps, args.warmup, stream)
        cb.fluid.set_graph(False)
        stats = kernel_stats(root, cb.step, 2, stream)
        ends[precision] = cb.fluid.get_state("pos")
        entry = {"ms_per_step": ms, "value": out["particles"] / (ms / 1e3),
                 "kernels_ms_per_step": per_step(stats, 2),
                 "eroded_tiles_mip0_last_frame": int(sum(cb.editor.readTileCounts(s)[0] for s in range(args.scenes)))}
        if precision == args.precision:
            out.update(entry)
            out["live_voxels"] = int(sum(int(cb.fluid.live_tiles(s, m).sum()) * 512 for s in range(args.scenes)
                                         for m in range(out["mips"])))
            out["roofline"] = roofline_of(root, stats, out["particles"], out["mips"], out["live_voxels"], "coupled")
        else:
            out["plain_f32"] = entry
        cb.destroy()
    if "plain_f32" in out:
        d = np.linalg.norm(ends[args.precision].astype(np.float64) - ends["f32"], axis=1) / 0.1
        out["error"] = {"vs": f"f32 coupled frames, same start, after {args.warmup + args.steps + 2} frames",
                        "position_over_h_rms": float(np.sqrt(np.mean(d * d))), "position_over_h_max": float(d.max())}
        out["fp16lds_over_f32"] = out["plain_f32"]["ms_per_step"] / out["ms_per_step"]
    return out


WIRE_METRIC = "baked-volume texels packed + unpacked per second (1 MI355X)"


def wire_bytes(kernel: str, texels: int, tiles: int, stored: int) -> int:
    """Algorithmic HBM bytes of one launch of a stream kernel: every (distance,
    material) pair of every texel is 4 bytes in the planes, a stored tile's texel 8
    bytes in the stream, a tile's table word 4 bytes and its bitmap bit 1/8 byte."""
    meta = 4 * tiles + tiles // 8
    return {"wire_classify": 4 * texels + 4 * (tiles - stored) + tiles // 8,
            "wire_write": 512 * stored * (4 + 8) + meta,
            "wire_unpack": 4 * texels + 512 * stored * 8 + meta}[kernel]


def measure_wire(root, args, stream) -> dict:
    """pack and unpack of the 288^3 x 3 ridge world (the dam-break-1m collider)
    through the stream format, the stream buffer resident in HBM; unpack is given
    the header, so neither synchronises."""
    import numpy as np

    import clayhip
    from clayhip import _lib as L
    from clayhip import scenes

    vol = clayhip.SdfVolume(root, resolution=288, mipLevels=3, worldSize=scenes.WORLD, origin=scenes.ORIGIN,
                            band=4)
    vol.bakeBrushes([dict(b, material=float(i + 1)) for i, b in enumerate(scenes.ridge_world_brushes())])
    buf, length = vol.pack_device(stream=stream)
    hdr = L.WireHeader.from_buffer_copy(buf[:96].cpu().numpy().tobytes())
    texels = sum((288 >> m) ** 3 for m in range(3))
    tiles, stored = int(hdr.tiles), int(hdr.stored_tiles)
    pack = lambda s: vol.pack_device(buf, stream=s, sync=False)
    unpack = lambda s: vol.unpack_device(buf, length, stream=s, header=hdr)
    pack_ms = timed(pack, args.steps, args.warmup, stream)
    unpack_ms = timed(unpack, args.steps, args.warmup, stream)
    stats = kernel_stats(root, lambda s: (pack(s), unpack(s)), args.steps, stream)
    kernels = {}
    for name in ("wire_classify", "wire_write", "wire_unpack"):
        tot_ms, n = stats[name]
        avg_s = tot_ms / n / 1e3
        nb = wire_bytes(name, texels, tiles, stored)
        kernels[name] = {"avg_ms": round(avg_s * 1e3, 4), "bytes_per_launch": nb,
                         "achieved_GBs": round(nb / avg_s / 1e9, 1)}
    dom = max(kernels, key=lambda k: kernels[k]["avg_ms"])
    d = kernels[dom]
    vol.wireStatus()
    host = buf[:length].cpu().numpy().tobytes()
    vol.destroy()
    return {"pack_ms": pack_ms, "unpack_ms": unpack_ms, "texels": texels, "tiles": tiles, "stored": stored,
            "stream_bytes": length, "plane_bytes": 4 * texels, "kernels": kernels,
            "roofline": {"bound": "hbm", "achieved": d["achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(d["achieved_GBs"] / HBM_PEAK_GBS, 5),
                         "traffic": committed_traffic(dom, "wire-288"), "kernel": dom, "avg_ms": d["avg_ms"],
                         "bytes_per_launch": d["bytes_per_launch"]},
            "stream": host}


def cpu_baseline_wire(stream: bytes, budget_s: float) -> dict:
    """The numpy restatement unpacking then packing the same stream on the host."""
    import numpy as np

    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O

    t0 = time.perf_counter()
    runs = 0
    texels = 0
    while True:
        h, dist, mat = O.wire_unpack(stream)
        O.wire_pack(dist, mat, float(h["band"]), float(h["world_size"]), tuple(h["origin"]),
                    float(h["gradient_voxels"]))
        texels += sum(x.size for x in dist)
        runs += 1
        if time.perf_counter() - t0 > budget_s or runs >= 20:
            break
    dt = time.perf_counter() - t0
    return {"value": texels / dt, "unit": "texels/s", "cores": 1, "kind": "port",
            "sample": f"{runs} unpack + pack round trips of the 288^3 x 3 world stream with the numpy restatement"}


def main_wire(root, args, stream) -> int:
    import torch

    t_wall = time.perf_counter()
    r = measure_wire(root, args, stream)
    ms = r["pack_ms"] + r["unpack_ms"]
    line = {
        "metric": WIRE_METRIC,
        "value": r["texels"] / (ms / 1e3),
        "unit": "texels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u16",
        "data": "synthetic (the ridge world baked at 288^3 x 3 mips)",
        "config": {"workload": "wire-288: pack + unpack of one 288^3 x 3-mip baked world",
                   "texels": r["texels"], "tiles": r["tiles"], "stored_tiles": r["stored"],
                   "stream_bytes": r["stream_bytes"], "plane_bytes": r["plane_bytes"],
                   "pack_ms": round(r["pack_ms"], 4), "unpack_ms": round(r["unpack_ms"], 4)},
        "roofline": r["roofline"],
        "kernels": r["kernels"],
    }
    line["wall_s"] = round(time.perf_counter() - t_wall, 3)
    line["cpu_baseline"] = None if args.no_cpu else cpu_baseline_wire(r["stream"], args.cpu_budget)
    root.destroy()
    print(json.dumps(line), flush=True)
    return 0


SWEEP_METRIC = "tracer sweeps (primary cone rays) per second over a baked world (1 MI355X)"


def camera_rays(width: int, height: int, fov_y: float, eye, at):
    """Primary cone rays of a pinhole camera, 10 floats each (ro, rd, tMin 0, tMax 64,
    radius 0, aperture = pixelAperture, march.ts:49-51), directions normalised in
    f32, ordered by 8 x 8 pixel tiles so a wave's rays are neighbours."""
    import numpy as np

    eye = np.asarray(eye, np.float64)
    fwd = np.asarray(at, np.float64) - eye
    fwd /= np.linalg.norm(fwd)
    right = np.cross(fwd, (0.0, 1.0, 0.0))
    right /= np.linalg.norm(right)
    up = np.cross(right, fwd)
    th = np.tan(fov_y / 2)
    ys, xs = np.mgrid[0:height, 0:width]
    ys = ys.reshape(height // 8, 8, width // 8, 8).transpose(0, 2, 1, 3).reshape(-1)
    xs = xs.reshape(height // 8, 8, width // 8, 8).transpose(0, 2, 1, 3).reshape(-1)
    u = ((xs + 0.5) / width * 2 - 1) * th * width / height
    v = (1 - (ys + 0.5) / height * 2) * th
    d = (fwd[None] + u[:, None] * right[None] + v[:, None] * up[None]).astype(np.float32)
    d /= np.sqrt((d * d).sum(1, dtype=np.float32))[:, None]
    rays = np.zeros((d.shape[0], 10), np.float32)
    rays[:, 0:3] = eye
    rays[:, 3:6] = d
    rays[:, 7] = 64.0
    rays[:, 9] = np.float32(th / (0.5 * height))
    return rays


def measure_sweep(root, args, stream) -> dict:
    """makeTracer(volumeField(world)).sweep of a 1920 x 1080 camera's primary cone
    rays over the 288^3 x 3 ridge world (the dam-break-1m collider), rays and hits
    resident in HBM."""
    import numpy as np
    import torch

    import clayhip
    from clayhip import scenes

    vol = clayhip.SdfVolume(root, resolution=288, mipLevels=3, worldSize=scenes.WORLD, origin=scenes.ORIGIN,
                            band=4)
    vol.bakeBrushes([dict(b, material=float(i + 1)) for i, b in enumerate(scenes.ridge_world_brushes())])
    field = clayhip.volumeField(vol)
    rays_h = camera_rays(1920, 1080, np.radians(60.0), (0.0, 6.0, -16.0), (0.0, 0.0, 0.0))
    n = rays_h.shape[0]
    rays = torch.from_numpy(rays_h).cuda()
    hits = torch.zeros((n, 4), dtype=torch.int32, device="cuda")
    step = lambda s: field.sweep_device(rays.data_ptr(), n, hits.data_ptr(), s.cuda_stream)
    ms = timed(step, args.steps, args.warmup, stream)
    stats = kernel_stats(root, step, args.steps, stream)
    h = hits.cpu().numpy()
    samples = int(h[:, 2].astype(np.int64).sum())
    tot_ms, launches = stats["field_sweep"]
    avg_s = tot_ms / launches / 1e3
    # Compulsory bytes: per ray the ray (40 B) and its hit (16 B).  The texels each
    # sample reads come through the caches (the counted fetch is the rays plus a
    # fraction of the world's mip-0 plane), so the kernel's bound is instruction issue.
    nb = 56 * n
    levels = [(vol.download(m)[..., 0], vol.download(m)[..., 1]) for m in range(3)]
    field.destroy()
    vol.destroy()
    roof = {"bound": "hbm", "achieved": round(nb / avg_s / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(nb / avg_s / 1e9 / HBM_PEAK_GBS, 5),
            "traffic": committed_traffic("field_sweep", "sweep-288"), "kernel": "field_sweep",
            "avg_ms": round(avg_s * 1e3, 4), "bytes_per_launch": nb, "samples_per_launch": samples}
    valu = valu_object(root, committed_valu("field_sweep", "sweep-288"), avg_s)
    if valu is not None:
        roof["valu"] = valu
    return {"ms": ms, "rays": n, "samples": samples, "hit_frac": float((h[:, 1] != 0).mean()),
            "steps_mean": samples / n, "steps_max": int(h[:, 2].max()), "avg_ms": avg_s * 1e3, "roofline": roof,
            "rays_host": rays_h, "levels": levels}


def cpu_baseline_sweep(r: dict, budget_s: float) -> dict:
    """The oracle's f32 sweep over every 64th ray of the same camera and world."""
    import numpy as np

    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O
    from clayhip import scenes

    lv = r["levels"]
    col = O.Collider().add_volume(lv[0][0], 288, scenes.ORIGIN, scenes.WORLD, 4.0, mat16=lv[0][1], levels=lv)
    threads = min(16, os.cpu_count() or 1)
    O.set_threads(threads)
    sample = r["rays_host"][::64]
    t0 = time.perf_counter()
    runs = 0
    while True:
        col.sweep(sample)
        runs += 1
        if time.perf_counter() - t0 > budget_s or runs >= 20:
            break
    dt = time.perf_counter() - t0
    return {"value": runs * sample.shape[0] / dt, "unit": "rays/s", "cores": threads, "kind": "port",
            "sample": f"{runs} passes of every 64th ray ({sample.shape[0]}) with the oracle's f32 sweep, "
                      f"OpenMP over {threads} threads"}


def main_sweep(root, args, stream) -> int:
    t_wall = time.perf_counter()
    r = measure_sweep(root, args, stream)
    line = {
        "metric": SWEEP_METRIC,
        "value": r["rays"] / (r["ms"] / 1e3),
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (the ridge world baked at 288^3 x 3 mips, a 1920 x 1080 camera)",
        "config": {"workload": "sweep-288: 1920 x 1080 primary cone rays over one 288^3 x 3-mip world",
                   "rays": r["rays"], "samples": r["samples"], "samples_per_s": r["samples"] / (r["ms"] / 1e3),
                   "steps_mean": round(r["steps_mean"], 3), "steps_max": r["steps_max"],
                   "hit_frac": round(r["hit_frac"], 4)},
        "roofline": r["roofline"],
    }
    line["wall_s"] = round(time.perf_counter() - t_wall, 3)
    line["cpu_baseline"] = None if args.no_cpu else cpu_baseline_sweep(r, args.cpu_budget)
    root.destroy()
    print(json.dumps(line), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="batched-256",
                    choices=["batched-256", "dam-break-1m", "dam-break-16k", "coupled", "wire", "sweep"])
    ap.add_argument("--scenes", type=int, default=8, help="coupled workload: scenes of 524 288 + 4096 particles")
    ap.add_argument("--splat", default="tiled", choices=["tiled", "plain"], help="single-scene workloads only")
    ap.add_argument("--precision", default=None, choices=["f32", "fp16lds"],
                    help="fluid kernels (default: fp16lds for the batch, f32 for a single scene)")
    ap.add_argument("--no-graph", action="store_true", help="batch: eager frames instead of a captured graph")
    ap.add_argument("--autotune", action="store_true", help="re-tune launch shapes (else the cache)")
    ap.add_argument("--no-compare", action="store_true", help="batch: skip the plain comparisons")
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
    batched = args.workload == "batched-256"
    coupled = args.workload == "coupled"
    if args.precision is None:
        args.precision = "fp16lds" if (batched or coupled) else "f32"

    import torch

    import clayhip

    root = clayhip.Root(0)
    stream = torch.cuda.Stream()
    if args.workload == "wire":
        return main_wire(root, args, stream)
    if args.workload == "sweep":
        return main_sweep(root, args, stream)
    t_wall = time.perf_counter()
    comparisons = None
    if batched:
        r = measure_batched(root, args, stream)
        if not args.no_compare:
            single = measure_single(root, "dam-break-1m", "tiled", "f32", args.steps, args.warmup, stream)
            comparisons = {
                "plain_per_scene": r["plain_per_scene"],
                "plain_single_1m": {"path": "one f32 FluidSim, " + single["workload"],
                                    "ms_per_step": single["ms_per_step"], "value": single["value"]},
                "batched_over_plain_per_scene": r["value"] / r["plain_per_scene"]["value"],
                "batched_over_plain_single_1m": r["value"] / single["value"],
                "batched_1m_total": r["batched_1m"],
                "batched_1m_total_over_plain_single_1m": r["batched_1m"]["value"] / single["value"],
            }
    elif coupled:
        r = measure_coupled(root, args, stream)
    else:
        r = measure_single(root, args.workload, args.splat, args.precision, args.steps, args.warmup, stream,
                           autotune_now=args.autotune)
    wall = time.perf_counter() - t_wall
    scene = r["scene"]
    line = {
        "metric": METRIC,
        "value": r["value"] * 1.0,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == "f32" else "f32 (neighbour offsets staged as f16 in LDS)",
        "data": "synthetic (seeded jittered lattices)",
        "config": {
            "workload": r["workload"],
            "particles": r["particles"],
            "substeps": r["substeps"],
            "bake": (f"{scene['options']['bakeResolution']}^3 x {r['mips']} mips"
                     + (" following, per scene" if batched else "")),
            "precision": args.precision,
            "live_voxels": r["live_voxels"],
        },
        "step_roofline": step_roofline(r["particles"], r["substeps"], r["mips"], r["live_voxels"], r["ms_per_step"]),
        "roofline": r["roofline"],
        "splat_atomics": r.get("splat_atomics"),
        "kernels_ms_per_step": r["kernels_ms_per_step"],
    }
    if batched:
        line["config"].update({"scenes": r["scenes"], "graph": r["graph"], "tuning": r["tuning"]})
        line["error"] = r["error"]
        line["comparisons"] = comparisons
    elif coupled:
        line["config"].update({"scenes": r["scenes"], "bake": "64^3 x 3 mips following + 32^3 x 2 body + 288^3 x 3 world edits"})
        line["error"] = r.get("error")
        line["comparisons"] = {"plain_f32": r.get("plain_f32"), "fp16lds_over_f32": r.get("fp16lds_over_f32"),
                               "eroded_tiles_mip0_last_frame": r.get("eroded_tiles_mip0_last_frame")}
    else:
        line["config"]["splat"] = args.splat
        line["config"]["tuning"] = r["tuning"]
    line["wall_s"] = round(wall, 3)
    if not args.no_cpu and not coupled:
        threads = min(16, os.cpu_count() or 1)
        line["cpu_baseline"] = (cpu_baseline_batched(scene, args.cpu_budget, threads) if batched
                                else cpu_baseline_single(scene, args.cpu_budget, threads))
    else:
        line["cpu_baseline"] = None
    root.destroy()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
