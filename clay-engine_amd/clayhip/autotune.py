"""Launch-shape autotuner for a FluidBatch (or FluidSim).

Tries each candidate clay_tuning on the live workload -- a few captured frames
per candidate, timed with HIP events on one stream -- keeps the fastest, and
caches the choice in a JSON file keyed by device and workload.  The knobs only
change launch shapes (workgroup size, LDS staging slice, splat work-item size,
persistent grid, XCD-contiguous workgroup numbering) and, for the f32 path, whether
neighbour rows are staged in LDS (exact_lds: the same sums in the same order), never
results.  The cell size is not a knob: it would change every neighbour's summation
order, and a finer grid costs more per-row work than it saves in candidates.
"""
from __future__ import annotations

import itertools
import json
import os
import sys
from pathlib import Path

DEFAULT_CACHE = Path(__file__).resolve().parent / "tuned.json"

SPACE = {
    "block": (128, 256),
    "stage": (64, 128, 256),
    "chunk": (1024, 2048, 4096),
    "groups_per_cu": (4, 8, 16, 32),
    "fluid_xcd": (0, 1),
    "splat_xcd": (0, 1),
    "splat_owner": (0, 1),
    "early_bin": (0, 1),
}


def space_for(precision: str) -> dict:
    """The search space of a workload: f32 also tries the global-memory neighbour
    walk against the LDS-staged one; FP16-LDS always stages."""
    return dict(SPACE, exact_lds=(0, 1)) if precision == "f32" else dict(SPACE)


def candidates(space: dict = SPACE) -> list[dict]:
    keys = list(space)
    return [dict(zip(keys, vals)) for vals in itertools.product(*(space[k] for k in keys))]


def _device_name() -> str:
    """Architecture and compute-unit count: stable across drivers and tools,
    unlike the marketing name."""
    import torch

    props = torch.cuda.get_device_properties(0)
    arch = getattr(props, "gcnArchName", "") or torch.cuda.get_device_name(0)
    return f"{arch.split(':')[0]}x{props.multi_processor_count}"


def time_frames(obj, frames: int, stream) -> float:
    """Milliseconds per step of `obj` over `frames` steps on `stream` (one warm step first)."""
    import torch

    obj.step(stream)
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(frames):
        obj.step(stream)
    stop.record(stream)
    stop.synchronize()
    return start.elapsed_time(stop) / frames


def autotune(obj, key: str, frames: int = 3, space: dict = SPACE, cache: Path | None = DEFAULT_CACHE,
             state=None) -> dict:
    """Tunes `obj` in place and returns {"tuning", "ms", "table"}.  `state`, if
    given, is a callable that restores the workload's initial state before each
    candidate, so every candidate times the same frames.  The cached "ms" is this
    comparison's own figure (the first frames after the initial state), not a
    steady-state frame time."""
    import torch

    stream = torch.cuda.Stream()
    table = []
    cands = candidates(space)
    for i, cand in enumerate(cands):
        if state is not None:
            state()
        obj.tuning = cand
        table.append({"tuning": cand, "ms": time_frames(obj, frames, stream)})
        if (i + 1) % 32 == 0 or i + 1 == len(cands):
            best = min(table, key=lambda r: r["ms"])
            print(f"autotune {key}: {i + 1}/{len(cands)} candidates, best {best['ms']:.3f} ms",
                  file=sys.stderr, flush=True)
    best = min(table, key=lambda r: r["ms"])
    obj.tuning = best["tuning"]
    if state is not None:
        state()
    result = {"tuning": best["tuning"], "ms": best["ms"], "table": table}
    if cache is not None:
        store(cache, key, {"tuning": best["tuning"], "ms": best["ms"]})
    return result


def cache_key(workload: str, precision: str) -> str:
    return f"{_device_name()}|{workload}|{precision}"


def load(cache: Path, key: str) -> dict | None:
    try:
        data = json.loads(Path(cache).read_text())
    except (OSError, ValueError):
        return None
    entry = data.get(key)
    return entry.get("tuning") if isinstance(entry, dict) else None


def store(cache: Path, key: str, entry: dict) -> None:
    cache = Path(cache)
    try:
        data = json.loads(cache.read_text())
    except (OSError, ValueError):
        data = {}
    data[key] = entry
    tmp = cache.with_suffix(".tmp")
    tmp.write_text(json.dumps(data, indent=1, sort_keys=True) + "\n")
    os.replace(tmp, cache)
