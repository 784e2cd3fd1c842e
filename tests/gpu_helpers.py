"""Builds a GPU FluidSim and the matching CPU oracle state from one workload."""
from __future__ import annotations

import numpy as np

import clayhip
import oracle as O
from clayhip import make_brush


def bits_equal(a: np.ndarray, b: np.ndarray) -> bool:
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def describe_diff(name: str, got: np.ndarray, want: np.ndarray) -> str:
    got = np.asarray(got, dtype=np.float64)
    want = np.asarray(want, dtype=np.float64)
    bad = got != want
    if not bad.any():
        return f"{name}: equal"
    idx = np.argwhere(bad)[:5]
    return (f"{name}: {int(bad.sum())} of {bad.size} differ, max |d| = {np.nanmax(np.abs(got - want)):.3g}; "
            f"first at {idx.tolist()}")


def make_world(root, scene):
    spec = scene["collider"]
    if spec["kind"] == "analytic":
        world = clayhip.analyticField(root, spec["brushes"])
        col = O.Collider().add_analytic([make_brush(b) for b in spec["brushes"]])
        return world, col, None
    v = spec["volume"]
    vol = clayhip.SdfVolume(root, v)
    vol.bakeBrushes(spec["brushes"])
    world = clayhip.volumeField(vol)
    tex0 = vol.download(0)
    col = O.Collider().add_volume(tex0[..., 0], v["resolution"], v["origin"], v["worldSize"], v["band"])
    return world, col, vol


def make_pair(root, scene, splat_mode: str = "plain", precision: str = "f32", pos=None):
    opts = dict(scene["options"])
    opts["splatMode"] = splat_mode
    world, col, vol = make_world(root, scene)
    pos = scene["pos"] if pos is None else pos
    sim = clayhip.FluidSim(root, world, opts)
    sim.set_state(pos)
    orc = O.Fluid(opts, pos, collider=col, precision=precision)
    return sim, orc, (world, vol)


def surface_opts(scene) -> O.OrcSplatOpts:
    """Oracle splat options of a workload's FluidSim surface."""
    o = scene["options"]
    return O.splat_opts(o["bakeResolution"], o["bakeMipLevels"], o["worldSize"], o["origin"],
                        o["bakeBand"], o["surfaceRadius"], o.get("material", 0.0))
