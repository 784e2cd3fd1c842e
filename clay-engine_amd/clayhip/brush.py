"""Brush records: the host half of BrushSet.make (lib/src/field/brush.ts:360-591).

Only the analytic kinds the collider slice needs are evaluated on the GPU
(sphere, box, roundBox, capsule, torus, cylinder, plane); asking for another
kind raises, as BrushSet.kindId does for an unknown one (brush.ts:513-523).
"""
from __future__ import annotations

import math
from typing import Iterable, Mapping

from ._lib import Brush

# BuiltinBrushKind (brush.ts:19-35)
KINDS = {
    "sphere": 0,
    "box": 1,
    "roundBox": 2,
    "capsule": 3,
    "torus": 4,
    "cylinder": 5,
    "plane": 6,
    "boxFrame": 7,
    "volume": 8,
    "cone": 9,
    "octahedron": 10,
    "tetrahedron": 11,
    "dodecahedron": 12,
    "icosahedron": 13,
}
SUPPORTED = {"sphere", "box", "roundBox", "capsule", "torus", "cylinder", "plane"}
# BrushOp (brush.ts:41-48)
OPS = {"add": 0, "cut": 1, "paint": 2}


def euler_to_quat(x: float, y: float, z: float) -> tuple[float, float, float, float]:
    """brush.ts:150-161 (X then Y then Z)."""
    cx, sx = math.cos(x / 2), math.sin(x / 2)
    cy, sy = math.cos(y / 2), math.sin(y / 2)
    cz, sz = math.cos(z / 2), math.sin(z / 2)
    return (
        sx * cy * cz + cx * sy * sz,
        cx * sy * cz - sx * cy * sz,
        cx * cy * sz + sx * sy * cz,
        cx * cy * cz - sx * sy * sz,
    )


def quat_mul(a, b) -> tuple[float, float, float, float]:
    """Hamilton product, xyzw; `a` applied after `b` (brush.ts:164-174)."""
    return (
        a[3] * b[0] + a[0] * b[3] + a[1] * b[2] - a[2] * b[1],
        a[3] * b[1] - a[0] * b[2] + a[1] * b[3] + a[2] * b[0],
        a[3] * b[2] + a[0] * b[1] - a[1] * b[0] + a[2] * b[3],
        a[3] * b[3] - a[0] * b[0] - a[1] * b[1] - a[2] * b[2],
    )


def normalize_size(size) -> tuple[float, float, float]:
    """brush.ts:176-184."""
    if size is None:
        return (1.0, 1.0, 1.0)
    if isinstance(size, (int, float)):
        return (float(size),) * 3
    return (float(size[0]), float(size[1]), float(size[2]))


def builtin_bound(kind: str, size, radius: float) -> float:
    """Conservative influence radius at unit scale (brush.ts:187-221)."""
    sx, sy, sz = size
    if kind == "sphere":
        return sx
    if kind in ("box", "boxFrame", "roundBox"):
        return math.hypot(sx, sy, sz) + radius
    if kind == "capsule":
        return sy + radius
    if kind == "torus":
        return sx + radius
    if kind == "cylinder":
        return math.hypot(sx, sy)
    if kind == "cone":
        return math.hypot(max(sx, sz), sy)
    if kind in ("octahedron", "tetrahedron", "dodecahedron", "icosahedron"):
        return sx
    if kind == "plane":
        return math.inf
    if kind == "volume":
        return sx * math.sqrt(3)
    raise ValueError(kind)


def kind_id(kind: str) -> int:
    if kind not in KINDS:
        known = ", ".join(KINDS)
        raise ValueError(
            f"BrushSet: unknown primitive '{kind}'. Register it in `Game.create({{ brushes }})`. "
            f"Known: {known}"
        )
    return KINDS[kind]


def make_brush(desc: Mapping) -> Brush:
    """BrushSet.make (brush.ts:539-564): a BrushDesc dict -> 80-byte Brush record."""
    kind = desc["kind"]
    kid = kind_id(kind)
    if kind not in SUPPORTED:
        raise ValueError(
            f"brush kind '{kind}' is outside the GPU collider slice "
            f"({', '.join(sorted(SUPPORTED))})"
        )
    size = normalize_size(desc.get("size"))
    scale = float(desc.get("scale", 1.0))
    radius = float(desc.get("radius", 0.0))
    smooth = float(desc.get("smooth", 0.02))
    ex, ey, ez = desc.get("euler", (0.0, 0.0, 0.0))
    bound = builtin_bound(kind, size, radius) * scale + smooth
    q = euler_to_quat(ex, ey, ez)
    if desc.get("quat") is not None:
        q = quat_mul(desc["quat"], q)
    b = Brush()
    pos = desc.get("pos", (0.0, 0.0, 0.0))
    for i in range(3):
        b.pos[i] = float(pos[i])
        b.size[i] = size[i]
    for i in range(4):
        b.rot[i] = q[i]
    b.scale = scale
    b.radius = radius
    b.material = float(desc.get("material", 0.0))
    b.blend = smooth
    b.bound = bound if math.isfinite(bound) else 1e9
    b.mask = float(desc.get("onlyMaterial", -1.0))
    b.kind = kid
    b.op = OPS[desc.get("op", "add")]
    b.slot = int(desc.get("slot", 0))
    b.pad = 0
    return b


def compile_brushes(descs: Iterable[Mapping]) -> list[Brush]:
    """BrushSet.compile (brush.ts:567-573): stable sort by layer, then make."""
    indexed = list(enumerate(descs))
    indexed.sort(key=lambda t: (t[1].get("layer", 0), t[0]))
    return [make_brush(d) for _, d in indexed]


def brush_array(brushes) -> "Brush * n":
    items = [b if isinstance(b, Brush) else make_brush(b) for b in brushes]
    arr = (Brush * len(items))()
    for i, b in enumerate(items):
        arr[i] = b
    return arr
