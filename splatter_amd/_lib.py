"""ctypes binding of include/splatter.h (libsplatter_amd.so, built in-tree).

Loading fails loudly when the library has not been built: there is no other
implementation behind this package.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsplatter_amd.so")


class SplConfiguration(C.Structure):
    """struct spl_configuration (include/splatter.h)."""

    _fields_ = [
        ("frame_format", C.c_int),
        ("depth_sorting", C.c_int),
        ("use_covariance_for_scale", C.c_int32),
        ("use_unaligned_rectangles", C.c_int32),
        ("spherical_harmonics_order", C.c_uint32),
        ("max_splat_count", C.c_uint64),
        ("radix_bits_per_digit", C.c_uint32),
        ("frustum_culling_tolerance", C.c_float),
        ("ellipse_margin", C.c_float),
        ("splat_scale", C.c_float),
        ("field_of_view_y", C.c_float),
        ("blend_mode", C.c_int),
        ("tile_pairs_per_splat", C.c_uint32),
        ("max_views", C.c_uint32),
        ("use_hip_graphs", C.c_int32),
        ("sort_tile_keys", C.c_uint32),
        ("raster_mode", C.c_uint32),
        ("raster_batch", C.c_uint32),
        ("raster_longest_first", C.c_uint32),
    ]


class SplServerConfiguration(C.Structure):
    """struct spl_server_configuration (include/splatter.h)."""

    _fields_ = [
        ("renderer", SplConfiguration),
        ("max_in_flight", C.c_uint32),
        ("max_width", C.c_uint32),
        ("max_height", C.c_uint32),
        ("memory_budget_bytes", C.c_uint64),
    ]


class SplRequestInfo(C.Structure):
    """struct spl_request_info (include/splatter.h)."""

    _fields_ = [
        ("status", C.c_int),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("batch_views", C.c_uint32),
        ("submit_ms", C.c_double),
        ("start_ms", C.c_double),
        ("done_ms", C.c_double),
    ]


class SplServerStats(C.Structure):
    """struct spl_server_stats (include/splatter.h)."""

    _fields_ = [
        ("submitted", C.c_uint64),
        ("completed", C.c_uint64),
        ("failed", C.c_uint64),
        ("batches", C.c_uint64),
        ("batched_views", C.c_uint64),
        ("pending", C.c_uint64),
        ("slots", C.c_uint32),
        ("slots_busy", C.c_uint32),
        ("frames", C.c_uint32),
        ("reserved", C.c_uint32),
        ("memory_in_use", C.c_uint64),
        ("memory_peak", C.c_uint64),
        ("memory_budget", C.c_uint64),
    ]


# Every exported entry point with its signature: (name, restype, argtypes).
_P = C.c_void_p
_PP = C.POINTER(C.c_void_p)
_F16 = C.POINTER(C.c_float)
SIGNATURES = [
    ("spl_last_error", C.c_char_p, []),
    ("spl_configuration_default", None, [C.POINTER(SplConfiguration)]),
    ("spl_device_open", C.c_int, [C.c_int, _PP]),
    ("spl_device_close", C.c_int, [_P]),
    ("spl_synchronize", C.c_int, [_P]),
    ("spl_renderer_create", C.c_int, [_P, C.POINTER(SplConfiguration), _PP]),
    ("spl_renderer_destroy", C.c_int, [_P]),
    ("spl_renderer_create_async", C.c_int, [_P, C.POINTER(SplConfiguration), _PP]),
    ("spl_renderer_synchronize", C.c_int, [_P]),
    ("spl_renderer_get_config", C.c_int, [_P, C.POINTER(SplConfiguration)]),
    ("spl_scene_create", C.c_int, [_P, C.c_uint64, _PP]),
    ("spl_scene_new_dev_test", C.c_int, [_P, _PP]),
    ("spl_scene_splat_count", C.c_int, [_P, C.POINTER(C.c_uint64)]),
    ("spl_scene_destroy", C.c_int, [_P]),
    ("spl_parse_file_header", C.c_int, [C.c_char_p, C.POINTER(C.c_uint16), C.POINTER(C.c_uint64)]),
    ("spl_scene_load_chunk", C.c_int, [_P, C.c_char_p, C.c_uint16, C.c_uint64, C.c_uint64]),
    ("spl_scene_load_records", C.c_int, [_P, C.c_uint64, C.c_uint64, _P]),
    ("spl_frame_create", C.c_int, [_P, C.c_uint32, C.c_uint32, _PP]),
    ("spl_frame_destroy", C.c_int, [_P]),
    ("spl_frame_read", C.c_int, [_P, _P, C.c_uint64]),
    ("spl_render_frame", C.c_int, [_P, _P, _P, _F16, _F16]),
    ("spl_render_views", C.c_int, [_P, _P, C.c_uint32, _PP, _F16, _F16]),
    ("spl_render_frame_static", C.c_int, [_P, C.POINTER(SplConfiguration), _P, _P, _F16, _F16]),
    ("spl_server_configuration_default", None, [C.POINTER(SplServerConfiguration)]),
    ("spl_server_create", C.c_int, [_P, C.POINTER(SplServerConfiguration), _PP]),
    ("spl_server_destroy", C.c_int, [_P]),
    ("spl_server_scene_create", C.c_int, [_P, C.c_uint64, _PP]),
    ("spl_server_submit", C.c_int, [_P, _P, C.c_uint32, C.c_uint32, _F16, _F16, C.POINTER(C.c_uint64)]),
    ("spl_server_poll", C.c_int, [_P, C.POINTER(C.c_uint64), C.c_uint32, C.POINTER(C.c_uint32)]),
    ("spl_server_wait", C.c_int, [_P, C.c_uint64]),
    ("spl_server_request_info", C.c_int, [_P, C.c_uint64, C.POINTER(SplRequestInfo)]),
    ("spl_server_read", C.c_int, [_P, C.c_uint64, _P, C.c_uint64]),
    ("spl_server_release", C.c_int, [_P, C.c_uint64]),
    ("spl_server_get_stats", C.c_int, [_P, C.POINTER(SplServerStats)]),
    ("spl_debug_select_view", C.c_int, [_P, C.c_uint32]),
    ("spl_debug_visible_count", C.c_int, [_P, C.POINTER(C.c_uint64)]),
    ("spl_debug_read_entries", C.c_int, [_P, _P, _P, C.c_uint64, C.POINTER(C.c_uint64)]),
    ("spl_debug_read_records", C.c_int, [_P, _P, C.c_uint64, C.POINTER(C.c_uint64)]),
    ("spl_debug_enable_conics", C.c_int, [_P, C.c_int]),
    ("spl_debug_read_conics", C.c_int, [_P, _P, C.c_uint64, C.POINTER(C.c_uint64)]),
    ("spl_debug_pair_count", C.c_int, [_P, C.POINTER(C.c_uint64)]),
    ("spl_debug_enable_tile_stats", C.c_int, [_P, C.c_int]),
    ("spl_debug_read_tile_stats", C.c_int, [_P, _P, C.c_uint64, C.POINTER(C.c_uint64)]),
    ("spl_debug_enable_timing", C.c_int, [_P, C.c_int]),
    ("spl_debug_stage_times", C.c_int, [_P, _F16, C.c_uint32]),
]

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m splatter_amd.build` "
                "(the HIP library is the only implementation)")
        l = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


class SplatterError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"splatter error {code}: {message}")
        self.code = code


def check(rc: int) -> None:
    if rc != 0:
        raise SplatterError(rc, lib().spl_last_error().decode(errors="replace"))
