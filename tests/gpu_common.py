"""Shared set-up of the GPU parity tests: one device per process, workloads
rendered through the C ABI and, on the same inputs, through the oracle."""
from __future__ import annotations

import math
import sys

import numpy as np

import oracle_bind as ob
import splatter_amd as sa
from splatter_amd.synth import Workload, look_at, load_scene, orbit_poses, sphere_pose

_device = None


def device() -> sa.Device:
    global _device
    if _device is None:
        _device = sa.Device(0)
    return _device


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def config(**kw) -> sa.Configuration:
    c = sa.Configuration(spherical_harmonics_order=3, max_splat_count=1 << 20)
    for k, v in kw.items():
        setattr(c, k, v)
    return c


class Case:
    """One workload + renderer + scene + pose, with the oracle's answer."""

    def __init__(self, workload: Workload, cfg: sa.Configuration, width: int, height: int, pose):
        self.workload, self.cfg, self.width, self.height = workload, cfg, width, height
        self.cam, self.view = pose
        cfg.max_splat_count = max(cfg.max_splat_count, workload.count, 1)
        self.renderer = sa.Renderer(device(), cfg)
        self.records = workload.all()
        self.scene = sa.Scene(self.renderer, workload.count)
        if workload.count:
            self.scene.load_records(0, self.records)
        self.frame = sa.Frame(device(), width, height)
        self.splats = ob.decode(self.records) if workload.count else np.zeros(0, ob.SPLAT_DTYPE)
        self.u = ob.uniforms(self.cam, self.view, width, height, fov=cfg.field_of_view_y,
                             tol=cfg.frustum_culling_tolerance, margin=cfg.ellipse_margin,
                             scale=cfg.splat_scale)

    def render(self):
        self.renderer.render_frame(self.frame, self.cam, self.view, self.scene)
        img = self.frame.read()
        if self.cfg.frame_format == sa.FrameFormat.BGRA8_UNORM:
            img = img[..., [2, 1, 0, 3]]
        return img

    def oracle(self, with_image=True):
        c = self.cfg
        if c.depth_sorting == sa.DepthSorting.NONE:
            idx = ob.gpu_predicate_cull(self.splats["center"], ob.vp_of(self.u), c.frustum_culling_tolerance)
            rec = np.zeros((len(idx), 10), np.float32)
            for i0 in range(0, len(idx), 4096):
                r, _ = ob.project(self.splats[idx[i0:i0 + 4096]], self.u, c.spherical_harmonics_order,
                                  c.use_covariance_for_scale)
                rec[i0:i0 + len(r)] = r
            img = ob.raster(rec, self.u, c.use_unaligned_rectangles, c.blend_mode == sa.BlendMode.F32) \
                if with_image else None
            return None, idx, rec, img
        return ob.render_frame(self.splats, self.u, c.spherical_harmonics_order, c.use_covariance_for_scale,
                               c.use_unaligned_rectangles, c.blend_mode == sa.BlendMode.F32, with_image)


def image_diff(a: np.ndarray, b: np.ndarray) -> int:
    return int(np.abs(a.astype(np.int32) - b.astype(np.int32)).max()) if a.size else 0


def normwise_rel(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.maximum(np.abs(b).max(axis=-1), 1e-300)
    return np.abs(a - b).max(axis=-1) / den
