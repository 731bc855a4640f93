"""ctypes declarations of the C ABI in include/gravlume.h.

The shared library is built in-tree (gravlume_amd/libgravlume_hip.so); loading
fails loudly when it is missing, so no call can silently fall back to a CPU
path.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgravlume_hip.so")

GL_OK = 0
GL_ERR_INVALID_ARGUMENT = 1
GL_ERR_DOMAIN = 2
GL_ERR_NOT_REPRESENTABLE = 3
GL_ERR_DEVICE = 4
GL_ERR_CAPACITY = 5
STATUS_NAMES = {0: "GL_OK", 1: "GL_ERR_INVALID_ARGUMENT", 2: "GL_ERR_DOMAIN",
                3: "GL_ERR_NOT_REPRESENTABLE", 4: "GL_ERR_DEVICE", 5: "GL_ERR_CAPACITY"}

TERM_ACTIVE = 0
TERM_HORIZON = 1
TERM_ESCAPE = 2
TERM_SINGULARITY = 3
TERM_STEP_EXHAUSTION = 4
TERM_NUMERICAL_FAILURE = 5
TERM_UNCERTAIN = 6
TERM_SURFACE = 7
TERM_REJECT_EXHAUSTION = 8
TERM_CERTIFIED_CAPTURE = 9

EVENT_SINGULARITY = 1
EVENT_HORIZON = 2
EVENT_SURFACE = 4
EVENT_ESCAPE = 8

FLAG_NON_FINITE = 1
FLAG_INVALID_RADICAND = 2
FLAG_INVALID_DENOMINATOR = 4
FLAG_DRIFT_BUDGET = 8
FLAG_TURNING_ORDER = 16
FLAG_TIED_EVENTS = 32
FLAG_CERTIFIED_CAPTURE = 64
FLAG_RECONSTRUCTED = 128

OPT_CAPTURE_CERTIFICATE = 1
OPT_F64_HORIZON_STEPS = 2
OPT_F64_DISK_ROOT = 4
OPT_ESCAPE_MAP = 8
OPT_NEAR_CRITICAL_REFINE = 16

STATUS_PROVISIONAL_SHIFT = 16
STATUS_NEAR_CRITICAL = 1 << 20
STATUS_REFINED = 1 << 21
STATUS_BRANCH_UNSETTLED = 1 << 22
STATUS_NO_ADJOINT = 1 << 23

EMISSION_BOLOMETRIC = 0
EMISSION_BLACKBODY = 1
TRANSPORT_VACUUM = 0
TRANSPORT_SLAB = 1

BLACKBODY_LUT_ENTRIES = 4097
BLACKBODY_LUT_MIN_LOG2_K = -8.0
BLACKBODY_LUT_PER_OCTAVE = 128.0


class SceneDesc(C.Structure):
    _fields_ = [
        ("mass", C.c_double), ("spin", C.c_double), ("charge", C.c_double),
        ("chart", C.c_int32), ("options", C.c_uint32),
        ("observer_txyz", C.c_double * 4), ("target_txyz", C.c_double * 4),
        ("up_hint", C.c_double * 3), ("frequency", C.c_double),
        ("width", C.c_uint32), ("height", C.c_uint32), ("vertical_fov", C.c_double),
        ("subpixel", C.c_double * 2), ("escape_radius", C.c_double),
        ("horizon_offset", C.c_double), ("has_surface", C.c_int32),
        ("surface_inner", C.c_double), ("surface_outer", C.c_double),
        ("emission_model", C.c_uint32), ("transport_model", C.c_uint32),
        ("intensity_at_six", C.c_double), ("temperature_at_six", C.c_double),
        ("optical_depth", C.c_double), ("slab_emission", C.c_double),
        ("slab_temperature", C.c_double), ("rtol", C.c_double), ("atol", C.c_double),
        ("initial_step", C.c_double), ("minimum_step", C.c_double),
        ("maximum_step", C.c_double), ("maximum_step_per_radius", C.c_double),
        ("step_safety", C.c_double), ("step_min_factor", C.c_double),
        ("step_max_factor", C.c_double),
        ("max_accepted_steps", C.c_uint32), ("max_consecutive_rejects", C.c_uint32),
        ("event_tie_tolerance", C.c_double), ("arming_band", C.c_double),
        ("drift_limit", C.c_double), ("reproject_interval", C.c_uint32),
        ("reserved", C.c_uint32), ("near_critical_factor", C.c_double),
        ("branch_factor", C.c_double), ("refine_rtol", C.c_double),
        ("refine_atol", C.c_double),
    ]


class TileRegionC(C.Structure):
    _fields_ = [("scene", C.c_uint32), ("tile_x", C.c_uint32), ("tile_y", C.c_uint32),
                ("tiles_x", C.c_uint32), ("tiles_y", C.c_uint32)]


class RayOutputsC(C.Structure):
    _fields_ = [("status", C.c_void_p), ("steps", C.c_void_p), ("rejected", C.c_void_p),
                ("observable", C.c_void_p), ("travel_time", C.c_void_p), ("affine", C.c_void_p),
                ("drift", C.c_void_p), ("branch", C.c_void_p), ("radiance", C.c_void_p),
                ("rgba16f", C.c_void_p), ("event_residual", C.c_void_p),
                ("turning_radius", C.c_void_p), ("azimuth_advance", C.c_void_p),
                ("count", C.c_size_t)]


WORK_IMAGE = 0
WORK_GEODESIC_BATCH = 1


class WorkItem(C.Structure):
    _fields_ = [("scene", C.c_uint32), ("kind", C.c_uint32), ("states8", C.c_void_p),
                ("directions", C.c_void_p), ("count", C.c_size_t), ("outputs", RayOutputsC)]


class PersistentConfig(C.Structure):
    _fields_ = [("block_size", C.c_uint32), ("blocks_per_cu", C.c_uint32),
                ("refill_interval", C.c_uint32), ("refill_threshold", C.c_uint32),
                ("waves_per_simd", C.c_uint32)]


class QueueStats(C.Structure):
    _fields_ = [("items", C.c_uint64), ("refills", C.c_uint64),
                ("attempted_steps", C.c_uint64), ("lane_slots", C.c_uint64)]


# Every symbol include/gravlume.h declares, with its ctypes signature.
_P = C.c_void_p
SIGNATURES = {
    "gl_trace_plan_create": (C.c_int, [C.POINTER(SceneDesc), C.c_size_t, _P, C.c_uint32,
                                       C.c_uint32, C.POINTER(_P)]),
    "gl_trace_plan_destroy": (C.c_int, [_P]),
    "gl_trace_plan_scene_rays": (C.c_int, [_P, C.c_uint32, C.POINTER(C.c_size_t)]),
    "gl_trace_encode": (C.c_int, [_P, C.POINTER(TileRegionC), C.POINTER(RayOutputsC), _P]),
    "gl_trace_encode_persistent": (C.c_int, [_P, _P, C.c_size_t, C.POINTER(RayOutputsC), _P]),
    "gl_trace_plan_set_persistent_config": (C.c_int, [_P, C.POINTER(PersistentConfig)]),
    "gl_trace_plan_get_persistent_config": (C.c_int, [_P, C.POINTER(PersistentConfig)]),
    "gl_trace_plan_set_f64_overlap": (C.c_int, [_P, C.c_int]),
    "gl_trace_plan_flush": (C.c_int, [_P, _P]),
    "gl_geodesic_batch_trace": (C.c_int, [_P, C.c_uint32, _P, _P, C.c_size_t,
                                          C.POINTER(RayOutputsC), _P]),
    "gl_geodesic_batch_trace_noalloc": (C.c_int, [_P, C.c_uint32, _P, _P, C.c_size_t,
                                                  C.POINTER(RayOutputsC), _P]),
    "gl_geodesic_batch_trace_f64": (C.c_int, [_P, C.c_uint32, _P, _P, C.c_size_t, C.c_double,
                                              C.c_double, C.POINTER(RayOutputsC), _P]),
    "gl_scene_validate": (C.c_int, [C.POINTER(SceneDesc)]),
    "gl_blackbody_log2_fractions": (C.c_int, [C.c_double, C.POINTER(C.c_double)]),
    "gl_blackbody_lut": (C.c_int, [C.POINTER(C.c_float), C.c_size_t]),
    "gl_trace_encode_work": (C.c_int, [_P, C.POINTER(WorkItem), C.c_size_t, _P]),
    "gl_trace_encode_accelerated": (C.c_int, [_P, C.c_uint32, C.POINTER(RayOutputsC), _P]),
    "gl_trace_refine_shadow": (C.c_int, [_P, C.c_uint32, C.POINTER(RayOutputsC), _P]),
    "gl_trace_plan_shadow_edges": (C.c_int, [_P, C.POINTER(C.c_uint64), _P]),
    "gl_trace_surface_footprint": (C.c_int, [_P, C.c_uint32, _P, C.c_size_t, _P, _P, _P]),
    "gl_geodesic_batch_trace_persistent": (C.c_int, [_P, C.c_uint32, _P, _P, C.c_size_t,
                                                     C.POINTER(PersistentConfig),
                                                     C.POINTER(RayOutputsC), _P]),
    "gl_trace_plan_queue_stats": (C.c_int, [_P, C.POINTER(QueueStats), _P]),
    "gl_trace_plan_enable_timing": (C.c_int, [_P, C.c_int]),
    "gl_trace_plan_kernel_ms": (C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                         C.POINTER(C.c_uint64)]),
    "gl_frame_graph_create": (C.c_int, [_P, _P, C.c_size_t, C.POINTER(RayOutputsC), _P,
                                        C.POINTER(_P)]),
    "gl_frame_graph_launch": (C.c_int, [_P, _P]),
    "gl_frame_graph_destroy": (C.c_int, [_P]),
    "gl_vjp_forward": (C.c_int, [_P, C.c_uint32, C.c_double, C.c_uint32,
                                 C.POINTER(RayOutputsC), _P, C.POINTER(_P)]),
    "gl_vjp_backward": (C.c_int, [_P, _P, C.POINTER(C.c_double), _P, _P]),
    "gl_vjp_backward_observables": (C.c_int, [_P, _P, _P, _P, C.POINTER(C.c_double), _P, _P]),
    "gl_vjp_surface_footprint": (C.c_int, [_P, C.POINTER(RayOutputsC), _P, _P, _P]),
    "gl_vjp_destroy": (C.c_int, [_P]),
    "gl_last_error": (C.c_char_p, []),
    "gl_version": (C.c_char_p, []),
    "gl_flop_constants": (C.c_int, [C.POINTER(C.c_double), C.POINTER(C.c_double)]),
}

_lib = None


def load() -> C.CDLL:
    """Loads libgravlume_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python gravlume_amd/build.py` "
                "(the trace has no CPU fallback)")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


class GravlumeError(RuntimeError):
    """A C-ABI call returned a non-OK gl_status."""

    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


def check(status: int) -> None:
    if status != GL_OK:
        raise GravlumeError(status, load().gl_last_error().decode(errors="replace"))
