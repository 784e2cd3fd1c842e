"""ctypes binding of libclay_hip.so (include/clay_hip.h).

The library is the product: there is no CPU fallback.  Loading fails loudly
when the in-tree build is missing.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent.parent
LIB_PATH = Path(os.environ.get("CLAY_HIP_LIB", PKG_DIR / "libclay_hip.so"))

MAX_MIPS = 16
MAX_LEAVES = 4
CONTACTS = 32

OK = 0
ERR_INVALID = -1
ERR_HIP = -2
ERR_NODEVICE = -3
ERR_STATE = -4

LIVE_ALL, LIVE_PREFIX, LIVE_RANGE = 0, 1, 2
SPLAT_PLAIN, SPLAT_TILED = 0, 1
STATE_POS, STATE_VEL, STATE_DENSITY, STATE_PRESSURE, STATE_NVEL = range(5)
DEBUG_KEYS, DEBUG_PERM = 0, 1


class ClayError(RuntimeError):
    """Raised with the library's message when a clay_* call fails."""

    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


class Brush(C.Structure):
    """`Brush` (lib/src/field/brush.ts:53-80), 80 bytes."""

    _fields_ = [
        ("pos", C.c_float * 3),
        ("scale", C.c_float),
        ("rot", C.c_float * 4),
        ("size", C.c_float * 3),
        ("radius", C.c_float),
        ("material", C.c_float),
        ("blend", C.c_float),
        ("bound", C.c_float),
        ("mask", C.c_float),
        ("kind", C.c_uint32),
        ("op", C.c_uint32),
        ("slot", C.c_uint32),
        ("pad", C.c_uint32),
    ]


class TracerOpts(C.Structure):
    """`clay_tracer_opts`: makeTracer's options (lib/src/trace/march.ts:22-40)."""

    _fields_ = [("max_steps", C.c_int32), ("min_step", C.c_float), ("hit_eps", C.c_float), ("min_mip", C.c_int32)]


class SweepHit(C.Structure):
    """`clay_sweep_hit`: SweepHit (lib/src/trace/march.ts:4-12), 16 bytes."""

    _fields_ = [("t", C.c_float), ("hit", C.c_uint32), ("steps", C.c_uint32), ("dist", C.c_float)]


class VolumeDesc(C.Structure):
    _fields_ = [
        ("resolution", C.c_int32),
        ("mip_levels", C.c_int32),
        ("world_size", C.c_double),
        ("band", C.c_double),
        ("origin", C.c_double * 3),
        ("gradient_voxels", C.c_double),
    ]


class WireHeader(C.Structure):
    _fields_ = [
        ("magic", C.c_uint32), ("version", C.c_uint32), ("resolution", C.c_uint32), ("mip_levels", C.c_uint32),
        ("band", C.c_float), ("world_size", C.c_float), ("gradient_voxels", C.c_float), ("tiles", C.c_uint32),
        ("origin", C.c_double * 3), ("stored_tiles", C.c_uint32), ("reserved", C.c_uint32),
        ("table_offset", C.c_uint64), ("bitmap_offset", C.c_uint64), ("payload_offset", C.c_uint64),
        ("total_bytes", C.c_uint64),
    ]


WIRE_MAGIC, WIRE_VERSION = 0x42564C43, 1


class VolumeInfo(C.Structure):
    _fields_ = [
        ("resolution", C.c_int32),
        ("mip_levels", C.c_int32),
        ("tile_res", C.c_int32 * MAX_MIPS),
        ("cell_offsets", C.c_int32 * MAX_MIPS),
        ("total_cells", C.c_int32),
        ("pad", C.c_int32),
        ("world_size", C.c_double),
        ("voxel_size", C.c_double),
        ("band", C.c_double),
        ("origin", C.c_double * 3),
    ]


class AnalyticOpts(C.Structure):
    _fields_ = [("band", C.c_double), ("epsilon", C.c_double), ("material", C.c_double)]


class CloudDesc(C.Structure):
    _fields_ = [
        ("capacity", C.c_uint32),
        ("live_rule", C.c_int32),
        ("live_first", C.c_uint32),
        ("live_count", C.c_uint32),
        ("pos", C.c_void_p * 3),
        ("nrm", C.c_void_p * 3),
    ]


class SplatDesc(C.Structure):
    _fields_ = [
        ("radius", C.c_double),
        ("material", C.c_double),
        ("resolution", C.c_int32),
        ("mip_levels", C.c_int32),
        ("band", C.c_double),
        ("origin", C.c_double * 3),
        ("world_size", C.c_double),
        ("mode", C.c_int32),
        ("keep_codes", C.c_int32),
    ]


class FluidDesc(C.Structure):
    _fields_ = [
        ("capacity", C.c_uint32),
        ("bucket", C.c_int32),
        ("spacing", C.c_double),
        ("smoothing", C.c_double),
        ("density", C.c_double),
        ("radius", C.c_double),
        ("surface_radius", C.c_double),
        ("material", C.c_double),
        ("bake_resolution", C.c_int32),
        ("bake_mip_levels", C.c_int32),
        ("bake_band", C.c_double),
        ("origin", C.c_double * 3),
        ("world_size", C.c_double),
        ("substeps", C.c_int32),
        ("splat_mode", C.c_int32),
        ("gravity", C.c_double * 3),
        ("stiffness", C.c_double),
        ("viscosity", C.c_double),
        ("damping", C.c_double),
        ("max_speed", C.c_double),
        ("precision", C.c_int32),
        ("pad", C.c_int32),
        ("follow_size", C.c_double),
    ]


PRECISION_F32, PRECISION_FP16LDS = 0, 1


class Tuning(C.Structure):
    _fields_ = [
        ("block", C.c_int32),
        ("stage", C.c_int32),
        ("chunk", C.c_int32),
        ("groups_per_cu", C.c_int32),
        ("fluid_xcd", C.c_int32),
        ("splat_xcd", C.c_int32),
        ("splat_owner", C.c_int32),
        ("exact_lds", C.c_int32),
        ("early_bin", C.c_int32),
    ]


class BatchDesc(C.Structure):
    _fields_ = [
        ("scenes", C.c_uint32),
        ("pad", C.c_uint32),
        ("counts", C.POINTER(C.c_uint32)),
        ("fluid", FluidDesc),
    ]


class FluidParams(C.Structure):
    _fields_ = [
        ("gravity", C.c_double * 3),
        ("dt", C.c_double),
        ("emit_pos", C.c_double * 3),
        ("emit_radius", C.c_double),
        ("emit_vel", C.c_double * 3),
        ("max_speed", C.c_double),
        ("stiffness", C.c_double),
        ("viscosity", C.c_double),
        ("damping", C.c_double),
        ("contact_depth", C.c_double),
    ]


class FluidConsts(C.Structure):
    _fields_ = [
        ("h", C.c_double),
        ("mass", C.c_double),
        ("poly6", C.c_double),
        ("spiky_grad", C.c_double),
        ("rest_density", C.c_double),
        ("dt", C.c_double),
        ("radius", C.c_double),
        ("contact_depth", C.c_double),
        ("grid_origin", C.c_double * 3),
        ("grid_res", C.c_int32),
        ("bucket", C.c_int32),
        ("substeps", C.c_int32),
        ("pad", C.c_int32),
        ("wall_lo", C.c_double * 3),
        ("wall_hi", C.c_double * 3),
    ]


class EditorDesc(C.Structure):
    _fields_ = [("max_edits", C.c_int32), ("tile_capacity", C.c_int32), ("safety", C.c_double)]


REDUCE_F32_TREE, REDUCE_F64 = 0, 1
QUERY_SAMPLE, QUERY_FIELD = 0, 1

class SolverDesc(C.Structure):
    _fields_ = [
        ("gravity", C.c_double * 3),
        ("dt", C.c_double),
        ("substeps", C.c_int32),
        ("rotation_iterations", C.c_int32),
        ("radius", C.c_double),
        ("friction", C.c_double),
        ("damping", C.c_double),
        ("max_speed", C.c_double),
    ]


class BodyDesc(C.Structure):
    _fields_ = [
        ("first", C.c_uint32),
        ("count", C.c_uint32),
        ("flags", C.c_uint32),
        ("stiffness", C.c_float),
        ("plasticity", C.c_float),
        ("pad", C.c_float),
    ]


class ExtractorDesc(C.Structure):
    _fields_ = [
        ("resolution", C.c_uint32),
        ("body", C.c_uint32),
        ("base", C.c_uint32),
        ("capacity", C.c_uint32),
        ("motion", C.c_int32),
        ("instance", C.c_uint32),
    ]


BODY_KINEMATIC, BODY_WEIGHTLESS = 1, 2
FREE_PARTICLE = 0xFFFFFFFF

STRUCTS = {
    "clay_brush": Brush,
    "clay_volume_desc": VolumeDesc,
    "clay_volume_info": VolumeInfo,
    "clay_analytic_opts": AnalyticOpts,
    "clay_tracer_opts": TracerOpts,
    "clay_sweep_hit": SweepHit,
    "clay_cloud_desc": CloudDesc,
    "clay_splat_desc": SplatDesc,
    "clay_fluid_desc": FluidDesc,
    "clay_fluid_params": FluidParams,
    "clay_fluid_consts": FluidConsts,
    "clay_tuning": Tuning,
    "clay_batch_desc": BatchDesc,
    "clay_editor_desc": EditorDesc,
    "clay_solver_desc": SolverDesc,
    "clay_body_desc": BodyDesc,
    "clay_wire_header": WireHeader,
    "clay_extractor_desc": ExtractorDesc,
}

P = C.c_void_p
PP = C.POINTER(C.c_void_p)
I32 = C.c_int32
U32 = C.c_uint32
DP = C.POINTER(C.c_double)

SIGNATURES = {
    "clay_last_error": ([], C.c_char_p),
    "clay_version": ([], C.c_int),
    "clay_ctx_create": ([C.c_int, PP], C.c_int),
    "clay_ctx_destroy": ([P], C.c_int),
    "clay_ctx_synchronize": ([P], C.c_int),
    "clay_ctx_set_timing": ([P, C.c_int], C.c_int),
    "clay_ctx_kernel_stats": ([P, C.c_int, C.c_char_p, C.c_int, DP, C.POINTER(C.c_uint64)], C.c_int),
    "clay_ctx_reset_stats": ([P], C.c_int),
    "clay_volume_derive": ([I32, I32, C.POINTER(VolumeInfo)], C.c_int),
    "clay_volume_create": ([P, C.POINTER(VolumeDesc), PP], C.c_int),
    "clay_volume_destroy": ([P], C.c_int),
    "clay_volume_get_info": ([P, C.POINTER(VolumeInfo)], C.c_int),
    "clay_volume_set_origin": ([P, DP, P], C.c_int),
    "clay_volume_set_origin_device": ([P, P, P], C.c_int),
    "clay_volume_voxel_size_at": ([P, I32, DP], C.c_int),
    "clay_volume_download": ([P, I32, P], C.c_int),
    "clay_volume_upload": ([P, I32, P], C.c_int),
    "clay_volume_pack_bound": ([P, C.POINTER(C.c_uint64)], C.c_int),
    "clay_volume_pack": ([P, U32, P, C.c_uint64, C.POINTER(C.c_uint64), P], C.c_int),
    "clay_wire_check": ([C.POINTER(WireHeader), P], C.c_int),
    "clay_volume_unpack": ([P, U32, P, C.c_uint64, C.POINTER(WireHeader), P], C.c_int),
    "clay_volume_wire_status": ([P, C.POINTER(C.c_uint32)], C.c_int),
    "clay_volume_bake_brushes": ([P, C.POINTER(Brush), I32, P], C.c_int),
    "clay_field_analytic": ([P, C.POINTER(Brush), I32, C.POINTER(AnalyticOpts), PP], C.c_int),
    "clay_field_volume": ([P, P, PP], C.c_int),
    "clay_field_union": ([P, P, PP], C.c_int),
    "clay_field_offset": ([P, P, PP], C.c_int),
    "clay_field_lerp": ([P, P, P, PP], C.c_int),
    "clay_field_destroy": ([P], C.c_int),
    "clay_field_max_mip": ([P, C.POINTER(I32)], C.c_int),
    "clay_field_eval": ([P, P, I32, P, P], C.c_int),
    "clay_field_query": ([P, I32, I32, P, I32, P], C.c_int),
    "clay_field_entry": ([P, P, I32, P], C.c_int),
    "clay_field_scale": ([P, I32, C.POINTER(C.c_float), C.POINTER(C.c_float)], C.c_int),
    "clay_tracer_opts_default": ([C.POINTER(TracerOpts)], None),
    "clay_field_sweep": ([P, C.POINTER(TracerOpts), P, I32, P], C.c_int),
    "clay_field_sweep_device": ([P, C.POINTER(TracerOpts), P, I32, P, P], C.c_int),
    "clay_splat_create": ([P, C.POINTER(CloudDesc), C.POINTER(SplatDesc), PP], C.c_int),
    "clay_splat_destroy": ([P], C.c_int),
    "clay_splat_set_live": ([P, U32, U32], C.c_int),
    "clay_splat_bake": ([P, P], C.c_int),
    "clay_splat_bake_static": ([P, P, P, P, P], C.c_int),
    "clay_atomic_peak": ([P, I32, U32, DP], C.c_int),
    "clay_valu_peak": ([P, I32, U32, DP], C.c_int),
    "clay_fluid_splat_atomics": ([P, P], C.c_int),
    "clay_batch_splat_atomics": ([P, P], C.c_int),
    "clay_splat_set_center": ([P, DP, P], C.c_int),
    "clay_splat_volume": ([P, PP], C.c_int),
    "clay_splat_field": ([P, PP], C.c_int),
    "clay_splat_grid_download": ([P, I32, P], C.c_int),
    "clay_splat_tiles_download": ([P, I32, P], C.c_int),
    "clay_fluid_desc_default": ([C.POINTER(FluidDesc)], None),
    "clay_fluid_create": ([P, P, C.POINTER(FluidDesc), PP], C.c_int),
    "clay_fluid_destroy": ([P], C.c_int),
    "clay_fluid_constants": ([P, C.POINTER(FluidConsts)], C.c_int),
    "clay_fluid_set_state": ([P, P, P, U32], C.c_int),
    "clay_fluid_get_state": ([P, I32, P], C.c_int),
    "clay_fluid_spawn": ([P, U32, DP, DP, C.c_double], C.c_int),
    "clay_fluid_configure": ([P, C.POINTER(FluidParams)], C.c_int),
    "clay_fluid_get_params": ([P, C.POINTER(FluidParams)], C.c_int),
    "clay_fluid_step": ([P, P], C.c_int),
    "clay_fluid_set_graph": ([P, C.c_int32], C.c_int),
    "clay_fluid_substep": ([P, P], C.c_int),
    "clay_fluid_read_contacts": ([P, P, C.POINTER(U32), C.POINTER(U32)], C.c_int),
    "clay_fluid_conserved": ([P, DP], C.c_int),
    "clay_fluid_particle_count": ([P, C.POINTER(U32)], C.c_int),
    "clay_fluid_surface": ([P, PP], C.c_int),
    "clay_fluid_debug_download": ([P, I32, P], C.c_int),
    "clay_fluid_set_tuning": ([P, C.POINTER(Tuning)], C.c_int),
    "clay_splat_set_tuning": ([P, C.POINTER(Tuning)], C.c_int),
    "clay_fluid_get_tuning": ([P, C.POINTER(Tuning)], C.c_int),
    "clay_batch_create": ([P, P, C.POINTER(BatchDesc), PP], C.c_int),
    "clay_batch_destroy": ([P], C.c_int),
    "clay_batch_set_state": ([P, P, P], C.c_int),
    "clay_batch_get_state": ([P, I32, P], C.c_int),
    "clay_batch_step": ([P, P], C.c_int),
    "clay_batch_set_graph": ([P, I32], C.c_int),
    "clay_batch_configure": ([P, C.POINTER(FluidParams)], C.c_int),
    "clay_batch_set_tuning": ([P, C.POINTER(Tuning)], C.c_int),
    "clay_batch_get_tuning": ([P, C.POINTER(Tuning)], C.c_int),
    "clay_batch_constants": ([P, C.POINTER(FluidConsts)], C.c_int),
    "clay_batch_scene": ([P, U32, C.POINTER(U32), C.POINTER(U32)], C.c_int),
    "clay_batch_read_contacts": ([P, U32, P, C.POINTER(U32), C.POINTER(U32)], C.c_int),
    "clay_batch_volume_download": ([P, U32, I32, P], C.c_int),
    "clay_batch_grid_download": ([P, U32, I32, P], C.c_int),
    "clay_batch_tiles_download": ([P, U32, I32, P], C.c_int),
    "clay_batch_origin": ([P, U32, C.POINTER(C.c_float)], C.c_int),
    "clay_batch_conserved": ([P, U32, DP], C.c_int),
    "clay_editor_create": ([P, C.POINTER(EditorDesc), PP], C.c_int),
    "clay_editor_destroy": ([P], C.c_int),
    "clay_editor_push": ([P, C.POINTER(Brush), I32, C.POINTER(I32)], C.c_int),
    "clay_editor_pending": ([P, C.POINTER(I32)], C.c_int),
    "clay_editor_flush": ([P, P, C.POINTER(I32)], C.c_int),
    "clay_editor_erode": ([P, P, C.c_double, C.c_double, C.c_double, P], C.c_int),
    "clay_editor_erode_scene": ([P, P, U32, C.c_double, C.c_double, C.c_double, P], C.c_int),
    "clay_editor_refresh_all": ([P, P], C.c_int),
    "clay_editor_tile_counts": ([P, P], C.c_int),
    "clay_editor_cells_download": ([P, P], C.c_int),
    "clay_body_centroids": ([P, P, P, P, U32, I32, P, P], C.c_int),
    "clay_body_centroids_device": ([P, P, P, P, U32, I32, P, P], C.c_int),
    "clay_editor_push_scene": ([P, U32, C.POINTER(Brush), I32, C.POINTER(I32)], C.c_int),
    "clay_editor_erode_batch": ([P, P, C.c_double, C.c_double, C.c_double, P], C.c_int),
    "clay_editor_tile_counts_scene": ([P, U32, P], C.c_int),
    "clay_editor_cells_download_scene": ([P, U32, P], C.c_int),
    "clay_volume_create_instances": ([P, C.POINTER(VolumeDesc), U32, PP], C.c_int),
    "clay_volume_replicate": ([P, P], C.c_int),
    "clay_volume_instance_download": ([P, U32, I32, P], C.c_int),
    "clay_splat_create_scenes": ([P, C.POINTER(CloudDesc), C.POINTER(SplatDesc), U32, P, I32, P, PP], C.c_int),
    "clay_splat_instance_download": ([P, U32, I32, P], C.c_int),
    "clay_solver_desc_default": ([C.POINTER(SolverDesc)], None),
    "clay_solver_create": ([P, P, C.POINTER(SolverDesc), U32, U32, PP], C.c_int),
    "clay_solver_destroy": ([P], C.c_int),
    "clay_solver_set_bodies": ([P, C.POINTER(BodyDesc), U32], C.c_int),
    "clay_solver_add_body": ([P, C.POINTER(BodyDesc), P], C.c_int),
    "clay_solver_get_owners": ([P, P], C.c_int),
    "clay_solver_get_slices": ([P, P], C.c_int),
    "clay_solver_set_particles": ([P, P, P, P, U32], C.c_int),
    "clay_solver_set_force": ([P, U32, DP, DP], C.c_int),
    "clay_solver_step": ([P, P], C.c_int),
    "clay_solver_get_particles": ([P, I32, P], C.c_int),
    "clay_solver_get_bodies": ([P, P], C.c_int),
    "clay_solver_positions": ([P, P], C.c_int),
    "clay_extractor_create": ([P, P, C.POINTER(ExtractorDesc), PP], C.c_int),
    "clay_extractor_destroy": ([P], C.c_int),
    "clay_extractor_set_region": ([P, DP, C.c_double, DP, I32], C.c_int),
    "clay_extractor_extract": ([P, P], C.c_int),
    "clay_extractor_read": ([P, P, P, P], C.c_int),
    "clay_extractor_quad_capacity": ([P, P], C.c_int),
    "clay_extractor_motion": ([P, P, P], C.c_int),
}

FOLLOW_NONE, FOLLOW_BOUNDS, FOLLOW_CENTERS = 0, 1, 2

_lib: C.CDLL | None = None


def lib() -> C.CDLL:
    """Loads the in-tree library once.  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ClayError(
                ERR_STATE,
                f"{LIB_PATH} is missing: build it with `python clay-engine_amd/build.py` "
                "(there is no CPU fallback for the HIP path)",
            )
        # One HIP runtime per process: torch ships its own libamdhip64.so.7 under the
        # same soname, and whichever copy is loaded first serves both.  Loading torch's
        # first keeps torch (streams, events, tensors) and this library on one runtime.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        handle = C.CDLL(str(LIB_PATH))
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.argtypes = args
            fn.restype = res
        _lib = handle
    return _lib


def check(rc: int) -> None:
    if rc != OK:
        msg = lib().clay_last_error()
        raise ClayError(rc, msg.decode() if msg else f"clay error {rc}")


def stream_handle(stream) -> int | None:
    """Accepts None, an int hipStream_t, or a torch.cuda.Stream."""
    if stream is None:
        return None
    if isinstance(stream, int):
        return stream
    handle = getattr(stream, "cuda_stream", None)
    if handle is None:
        raise TypeError(f"not a stream: {stream!r}")
    return int(handle)


def dvec(values) -> C.Array:
    arr = (C.c_double * 3)()
    for i in range(3):
        arr[i] = float(values[i])
    return arr
