"""ctypes binding of the CPU oracle (oracle/liboracle.so) for the tests.

The oracle is the checker: tests compare the HIP path against it; nothing in
splatter_amd imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle.so")

SPLAT_DTYPE = np.dtype([
    ("rotation", "<f4", 4), ("center", "<f4", 3), ("padding", "<f4"),
    ("scale", "<f4", 3), ("alpha", "<f4"), ("sh", "<f4", 48),
])
assert SPLAT_DTYPE.itemsize == 240


class Uniforms(C.Structure):
    _fields_ = [
        ("camera_matrix", C.c_float * 16), ("view_matrix", C.c_float * 16),
        ("view_projection_matrix", C.c_float * 16), ("view_size", C.c_float * 2),
        ("image_size", C.c_uint32 * 2), ("frustum_culling_tolerance", C.c_float),
        ("ellipse_size_bias", C.c_float), ("ellipse_margin", C.c_float), ("splat_scale", C.c_float),
    ]


assert C.sizeof(Uniforms) == 224

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_LIB):
            subprocess.run(["make", "-C", ORACLE_DIR], check=True, capture_output=True)
        l = C.CDLL(ORACLE_LIB)
        P = C.c_void_p
        l.orc_decode_serialized.argtypes = [P, C.c_size_t, P]
        l.orc_dev_test_scene.argtypes = [P]
        l.orc_perspective_projection.argtypes = [C.c_float, C.c_float, C.c_float, C.c_float, P]
        l.orc_mat4_multiplication.argtypes = [P, P, P]
        l.orc_camera_uniforms.argtypes = [P, P, C.c_uint32, C.c_uint32, C.c_float, C.c_float, C.c_float,
                                          C.c_float, C.POINTER(Uniforms)]
        l.orc_cpu_depth_sort.argtypes = [P, C.c_size_t, P, C.c_float, P, P]
        l.orc_cpu_depth_sort.restype = C.c_size_t
        l.orc_gpu_predicate_cull.argtypes = [P, C.c_size_t, P, C.c_float, P]
        l.orc_gpu_predicate_cull.restype = C.c_size_t
        l.orc_stable_sort_pairs.argtypes = [P, P, C.c_size_t, P]
        l.orc_project_f64.argtypes = [P, C.POINTER(Uniforms), C.c_int, C.c_int, P, P]
        l.orc_project_f32.argtypes = [P, C.POINTER(Uniforms), C.c_int, C.c_int, P, P]
        l.orc_contour_f64.argtypes = [P, P, P, C.POINTER(Uniforms), P]
        l.orc_sh_color.argtypes = [P, P, C.c_int, P]
        l.orc_sh_constant.argtypes = [C.c_int]
        l.orc_sh_constant.restype = C.c_float
        l.orc_exp2.argtypes = [C.c_float]
        l.orc_exp2.restype = C.c_float
        l.orc_raster.argtypes = [P, C.c_size_t, C.POINTER(Uniforms), C.c_int, C.c_int, P, C.c_int]
        l.orc_render_frame.argtypes = [P, C.c_size_t, C.POINTER(Uniforms), C.c_int, C.c_int, C.c_int,
                                       C.c_int, P, P, P, P]
        l.orc_render_frame.restype = C.c_size_t
        l.orc_time_cpu_depth_sort.argtypes = [P, C.c_size_t, P, C.c_float, C.c_int, C.POINTER(C.c_size_t)]
        l.orc_time_cpu_depth_sort.restype = C.c_double
        l.orc_cpu_depth_sort_threads.argtypes = [P, C.c_size_t, P, C.c_float, C.c_int, P, P]
        l.orc_cpu_depth_sort_threads.restype = C.c_size_t
        l.orc_time_cpu_depth_sort_threads.argtypes = [P, C.c_size_t, P, C.c_float, C.c_int, C.c_int,
                                                      C.POINTER(C.c_size_t)]
        l.orc_time_cpu_depth_sort_threads.restype = C.c_double
        _lib = l
    return _lib


def ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def decode(serialized: np.ndarray) -> np.ndarray:
    s = np.ascontiguousarray(serialized)
    out = np.zeros(len(s), SPLAT_DTYPE)
    lib().orc_decode_serialized(ptr(s), len(s), ptr(out))
    return out


def dev_test_scene() -> np.ndarray:
    out = np.zeros(3, SPLAT_DTYPE)
    lib().orc_dev_test_scene(ptr(out))
    return out


def uniforms(camera, view, width, height, fov=np.pi / 2, tol=1.1, margin=2.0, scale=1.0) -> Uniforms:
    u = Uniforms()
    cam = np.ascontiguousarray(np.asarray(camera, np.float32).reshape(16))
    vw = np.ascontiguousarray(np.asarray(view, np.float32).reshape(16))
    lib().orc_camera_uniforms(ptr(cam), ptr(vw), int(width), int(height), float(np.float32(fov)),
                              float(tol), float(margin), float(scale), C.byref(u))
    return u


def vp_of(u: Uniforms) -> np.ndarray:
    return np.array(u.view_projection_matrix[:], dtype=np.float32)


def cpu_depth_sort(positions: np.ndarray, vp: np.ndarray, tol: float):
    pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
    n = len(pos)
    keys = np.empty(max(n, 1), np.uint32)
    idx = np.empty(max(n, 1), np.uint32)
    vpa = np.ascontiguousarray(vp, dtype=np.float32).reshape(16)
    v = lib().orc_cpu_depth_sort(ptr(pos), n, ptr(vpa), float(tol), ptr(keys), ptr(idx))
    return keys[:v].copy(), idx[:v].copy()


def cpu_depth_sort_threads(positions: np.ndarray, vp: np.ndarray, tol: float, threads: int):
    """renderer.rs:372-396 on `threads` host threads (parallel transform and
    stable merge sort); the same entries as cpu_depth_sort."""
    pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
    n = len(pos)
    keys = np.empty(max(n, 1), np.uint32)
    idx = np.empty(max(n, 1), np.uint32)
    vpa = np.ascontiguousarray(vp, dtype=np.float32).reshape(16)
    v = lib().orc_cpu_depth_sort_threads(ptr(pos), n, ptr(vpa), float(tol), int(threads), ptr(keys), ptr(idx))
    return keys[:v].copy(), idx[:v].copy()


def gpu_predicate_cull(positions: np.ndarray, vp: np.ndarray, tol: float) -> np.ndarray:
    pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
    idx = np.empty(max(len(pos), 1), np.uint32)
    vpa = np.ascontiguousarray(vp, dtype=np.float32).reshape(16)
    v = lib().orc_gpu_predicate_cull(ptr(pos), len(pos), ptr(vpa), float(tol), ptr(idx))
    return idx[:v].copy()


def project(splats: np.ndarray, u: Uniforms, sh_order: int, covariance: bool, f32: bool = False):
    """Records (n, 10) f32 and conics (n, 6) for each splat."""
    s = np.ascontiguousarray(splats)
    rec = np.zeros((len(s), 10), np.float32)
    con = np.zeros((len(s), 6), np.float32 if f32 else np.float64)
    fn = lib().orc_project_f32 if f32 else lib().orc_project_f64
    for i in range(len(s)):
        fn(s[i:i + 1].ctypes.data_as(C.c_void_p), C.byref(u), int(sh_order), int(covariance),
           rec[i:i + 1].ctypes.data_as(C.c_void_p), con[i:i + 1].ctypes.data_as(C.c_void_p))
    return rec, con


def raster(records: np.ndarray, u: Uniforms, unaligned: bool = True, blend_f32: bool = False,
           brute_force: bool = False) -> np.ndarray:
    rec = np.ascontiguousarray(records, dtype=np.float32).reshape(-1, 10)
    w, h = u.image_size[0], u.image_size[1]
    img = np.zeros((h, w, 4), np.uint8)
    lib().orc_raster(ptr(rec), len(rec), C.byref(u), int(unaligned), int(blend_f32), ptr(img), int(brute_force))
    return img


def render_frame(splats: np.ndarray, u: Uniforms, sh_order: int, covariance: bool, unaligned: bool = True,
                 blend_f32: bool = False, with_image: bool = True):
    s = np.ascontiguousarray(splats)
    n = len(s)
    keys = np.empty(max(n, 1), np.uint32)
    idx = np.empty(max(n, 1), np.uint32)
    rec = np.empty((max(n, 1), 10), np.float32)
    img = np.zeros((u.image_size[1], u.image_size[0], 4), np.uint8) if with_image else None
    v = lib().orc_render_frame(ptr(s), n, C.byref(u), int(sh_order), int(covariance), int(unaligned),
                               int(blend_f32), ptr(keys), ptr(idx), ptr(rec),
                               ptr(img) if img is not None else None)
    return keys[:v].copy(), idx[:v].copy(), rec[:v].copy(), img


def exp2(y: float) -> float:
    return lib().orc_exp2(float(y))
