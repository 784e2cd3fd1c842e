"""clayhip -- MI355X (gfx950) fluid-to-field path of clay-engine, host side.

A Python mirror of the reference's TypeScript surface for this path
(FluidSim, SplatField, ParticleCloud, SdfVolume, TracedField and the field
combinators), calling hand-written HIP kernels through the C ABI in
include/clay_hip.h.  There is no CPU fallback: every call goes to the GPU.
"""
from ._lib import ClayError, LIB_PATH, lib
from .batch import FluidBatch
from .brush import compile_brushes, make_brush
from .edit import SdfEditor, body_centroids, body_centroids_device
from .field import TracedField, analyticField, lerpField, offsetField, unionField, volumeField
from .fluid import FluidSim, fluid_desc
from .pbd import ClaySolver, SurfaceExtractor
from .splat import ParticleCloud, SplatField, bake_static
from .volume import Root, SdfVolume, derive

__all__ = [
    "ClayError",
    "ClaySolver",
    "SurfaceExtractor",
    "FluidBatch",
    "FluidSim",
    "LIB_PATH",
    "ParticleCloud",
    "Root",
    "SdfEditor",
    "SdfVolume",
    "SplatField",
    "TracedField",
    "analyticField",
    "bake_static",
    "body_centroids",
    "body_centroids_device",
    "compile_brushes",
    "derive",
    "fluid_desc",
    "lib",
    "make_brush",
    "unionField",
    "offsetField",
    "lerpField",
    "volumeField",
]
