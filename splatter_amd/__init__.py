"""splatter_amd -- MI355X-native splat pipeline behind splatter's Renderer/Scene API.

Python mirror of the reference crate's public surface (src/lib.rs:1-3):

    renderer::DepthSorting      -> DepthSorting
    renderer::Configuration     -> Configuration
    renderer::Renderer::new     -> Renderer(device, config)
    Renderer::render_frame      -> Renderer.render_frame(frame, camera_matrix, view_matrix, scene)
    scene::Scene::new           -> Scene(renderer, splat_count)
    Scene::new_dev_test         -> Scene.new_dev_test(renderer)
    Scene::parse_file_header    -> Scene.parse_file_header(path)
    Scene::load_chunk           -> Scene.load_chunk(path, header_size, begin, end)
    Scene.splat_count           -> Scene.splat_count

wgpu's Device/Queue pair is a Device (one HIP device and stream) and the
TextureView + viewport_size pair is a Frame.  Everything runs in
libsplatter_amd.so (HIP kernels for gfx950); this module only marshals.
"""
from __future__ import annotations

import ctypes as C
import enum
import math
from dataclasses import dataclass, field

import numpy as np

from ._lib import (SplConfiguration, SplRequestInfo, SplServerConfiguration, SplServerStats, SplatterError,
                   check, lib)

__all__ = [
    "DepthSorting", "FrameFormat", "BlendMode", "Configuration", "Device", "Renderer",
    "Scene", "Frame", "Server", "SplatterError", "SERIALIZED_SPLAT_DTYPE",
]


class DepthSorting(enum.IntEnum):
    """renderer.rs:14-23"""
    NONE = 0
    CPU = 1
    GPU = 2
    GPU_INDIRECT_DRAW = 3


class FrameFormat(enum.IntEnum):
    BGRA8_UNORM = 0
    RGBA8_UNORM = 1


class BlendMode(enum.IntEnum):
    UNORM8 = 0
    F32 = 1


@dataclass
class Configuration:
    """renderer.rs:26-47 with the showcase's values (showcase.rs:39-50)."""
    frame_format: FrameFormat = FrameFormat.BGRA8_UNORM
    depth_sorting: DepthSorting = DepthSorting.GPU
    use_covariance_for_scale: bool = True
    use_unaligned_rectangles: bool = True
    spherical_harmonics_order: int = 2
    max_splat_count: int = 1024 * 1024 * 4
    radix_bits_per_digit: int = 8
    frustum_culling_tolerance: float = 1.1
    ellipse_margin: float = 2.0
    splat_scale: float = 1.0
    field_of_view_y: float = math.pi * 0.5
    blend_mode: BlendMode = BlendMode.UNORM8
    tile_pairs_per_splat: int = 4
    max_views: int = 1
    use_hip_graphs: bool = True
    sort_tile_keys: int = 4096
    raster_mode: int = 3             # per-lane walk over coverage masks (every mode renders the same image)
    raster_batch: int = 128          # batches of two or more views; one view rasters with 256
    raster_longest_first: int = 0    # batches of two or more views; one view always starts longest tiles first

    def to_c(self) -> SplConfiguration:
        c = SplConfiguration()
        c.frame_format = int(self.frame_format)
        c.depth_sorting = int(self.depth_sorting)
        c.use_covariance_for_scale = int(bool(self.use_covariance_for_scale))
        c.use_unaligned_rectangles = int(bool(self.use_unaligned_rectangles))
        c.spherical_harmonics_order = int(self.spherical_harmonics_order)
        c.max_splat_count = int(self.max_splat_count)
        c.radix_bits_per_digit = int(self.radix_bits_per_digit)
        c.frustum_culling_tolerance = float(self.frustum_culling_tolerance)
        c.ellipse_margin = float(self.ellipse_margin)
        c.splat_scale = float(self.splat_scale)
        c.field_of_view_y = float(np.float32(self.field_of_view_y))
        c.blend_mode = int(self.blend_mode)
        c.tile_pairs_per_splat = int(self.tile_pairs_per_splat)
        c.max_views = int(self.max_views)
        c.use_hip_graphs = int(bool(self.use_hip_graphs))
        c.sort_tile_keys = int(self.sort_tile_keys)
        c.raster_mode = int(self.raster_mode)
        c.raster_batch = int(self.raster_batch)
        c.raster_longest_first = int(self.raster_longest_first)
        return c


# scene.rs:17-24 SerializedSplat, 248 bytes
SERIALIZED_SPLAT_DTYPE = np.dtype([
    ("center", "<f4", 3), ("n", "<f4", 3), ("color", "<f4", 48),
    ("alpha", "<f4"), ("scale", "<f4", 3), ("rotation", "<f4", 4),
])
assert SERIALIZED_SPLAT_DTYPE.itemsize == 248


def _mat16(m) -> C.Array:
    a = np.ascontiguousarray(np.asarray(m, dtype=np.float32).reshape(16))
    return (C.c_float * 16)(*a.tolist())


class Device:
    """wgpu Device + Queue: one HIP device with one stream."""

    def __init__(self, ordinal: int = 0):
        h = C.c_void_p()
        check(lib().spl_device_open(int(ordinal), C.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def synchronize(self) -> None:
        check(lib().spl_synchronize(self._h))

    def render_frame_static(self, config: "Configuration", frame: "Frame", camera_matrix, view_matrix,
                            scene: "Scene") -> None:
        """Renderer.render_frame without a renderer and without allocating:
        the workspace is the library's fixed scratch array on this device
        (scenes up to 131072 splats, frames up to 3840x2160).  Calls on one
        device share that array and must not overlap."""
        c = config.to_c()
        check(lib().spl_render_frame_static(self._h, C.byref(c), scene.handle, frame.handle,
                                            _mat16(camera_matrix), _mat16(view_matrix)))

    def close(self) -> None:
        if self._h:
            lib().spl_device_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Frame:
    """Render target of width x height RGBA8/BGRA8 pixels in device memory."""

    def __init__(self, device: Device, width: int, height: int):
        h = C.c_void_p()
        check(lib().spl_frame_create(device.handle, int(width), int(height), C.byref(h)))
        self._h, self.device, self.width, self.height = h, device, int(width), int(height)

    @property
    def handle(self):
        return self._h

    def read(self) -> np.ndarray:
        out = np.empty((self.height, self.width, 4), dtype=np.uint8)
        check(lib().spl_frame_read(self._h, out.ctypes.data_as(C.c_void_p), out.nbytes))
        return out

    def close(self) -> None:
        if self._h:
            lib().spl_frame_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Renderer:
    """renderer.rs:64-84 Renderer."""

    def __init__(self, device: Device, config: Configuration | None = None, own_stream: bool = False):
        """own_stream: launches go to a stream of the renderer's own
        (spl_renderer_create_async), so the batches of two such renderers
        overlap on the device; call synchronize() before reading its frames."""
        self.config = config or Configuration()
        self.device = device
        h = C.c_void_p()
        cc = self.config.to_c()
        create = lib().spl_renderer_create_async if own_stream else lib().spl_renderer_create
        check(create(device.handle, C.byref(cc), C.byref(h)))
        self._h = h
        self.own_stream = own_stream

    def synchronize(self) -> None:
        """Waits for everything this renderer has queued."""
        check(lib().spl_renderer_synchronize(self._h))

    @property
    def handle(self):
        return self._h

    def render_frame(self, frame: Frame, camera_matrix, view_matrix, scene: "Scene") -> None:
        """renderer.rs:355-363: camera_matrix = motor3d_to_mat4(motor),
        view_matrix = motor3d_to_mat4(motor.inverse()), column-major 4x4."""
        check(lib().spl_render_frame(self._h, scene.handle, frame.handle,
                                     _mat16(camera_matrix), _mat16(view_matrix)))

    def render_views(self, frames, camera_matrices, view_matrices, scene: "Scene") -> None:
        n = len(frames)
        cams = np.ascontiguousarray(np.asarray(camera_matrices, dtype=np.float32).reshape(n, 16))
        views = np.ascontiguousarray(np.asarray(view_matrices, dtype=np.float32).reshape(n, 16))
        arr = (C.c_void_p * n)(*[f.handle.value for f in frames])
        check(lib().spl_render_views(self._h, scene.handle, n, C.cast(arr, C.POINTER(C.c_void_p)),
                                     cams.ctypes.data_as(C.POINTER(C.c_float)),
                                     views.ctypes.data_as(C.POINTER(C.c_float))))

    # ---- parity taps ---------------------------------------------------
    def select_view(self, view: int) -> None:
        """Debug reads describe this view of the last batch."""
        check(lib().spl_debug_select_view(self._h, int(view)))

    def visible_count(self) -> int:
        v = C.c_uint64()
        check(lib().spl_debug_visible_count(self._h, C.byref(v)))
        return int(v.value)

    def pair_count(self) -> int:
        v = C.c_uint64()
        check(lib().spl_debug_pair_count(self._h, C.byref(v)))
        return int(v.value)

    def read_entries(self):
        n = self.visible_count()
        keys = np.empty(max(n, 1), np.uint32)
        idx = np.empty(max(n, 1), np.uint32)
        cnt = C.c_uint64()
        check(lib().spl_debug_read_entries(self._h, keys.ctypes.data_as(C.c_void_p),
                                           idx.ctypes.data_as(C.c_void_p), n, C.byref(cnt)))
        return keys[:cnt.value], idx[:cnt.value]

    def read_records(self) -> np.ndarray:
        n = self.visible_count()
        out = np.empty((max(n, 1), 10), np.float32)
        cnt = C.c_uint64()
        check(lib().spl_debug_read_records(self._h, out.ctypes.data_as(C.c_void_p), n, C.byref(cnt)))
        return out[:cnt.value]

    def enable_conics(self, enable: bool = True) -> None:
        check(lib().spl_debug_enable_conics(self._h, int(enable)))

    def read_conics(self) -> np.ndarray:
        n = self.visible_count()
        out = np.empty((max(n, 1), 6), np.float64)
        cnt = C.c_uint64()
        check(lib().spl_debug_read_conics(self._h, out.ctypes.data_as(C.c_void_p), n, C.byref(cnt)))
        return out[:cnt.value]

    def enable_tile_stats(self, enable: bool = True) -> None:
        check(lib().spl_debug_enable_tile_stats(self._h, int(enable)))

    def tile_stats(self) -> np.ndarray:
        """(tiles, 2) array: list length and entries consumed per 16x16 tile."""
        cnt = C.c_uint64()
        check(lib().spl_debug_read_tile_stats(self._h, None, 0, C.byref(cnt)))
        out = np.zeros((max(cnt.value, 1), 2), np.uint32)
        check(lib().spl_debug_read_tile_stats(self._h, out.ctypes.data_as(C.c_void_p), cnt.value, C.byref(cnt)))
        return out[:cnt.value]

    def enable_timing(self, enable: bool = True) -> None:
        check(lib().spl_debug_enable_timing(self._h, int(enable)))

    STAGES = ("cull_key_compact", "digit_histograms", "depth_sort", "project", "bin", "pair_sort",
              "tile_ranges", "raster")

    def stage_times(self) -> np.ndarray:
        """Per-frame device milliseconds of each stage (Renderer.STAGES)."""
        out = (C.c_float * 8)()
        check(lib().spl_debug_stage_times(self._h, out, 8))
        return np.array(out[:], dtype=np.float64)

    def close(self) -> None:
        if self._h:
            lib().spl_renderer_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Scene:
    """scene.rs:35-41 Scene: splats resident in HBM."""

    def __init__(self, renderer: Renderer, splat_count: int, _handle=None):
        self.renderer = renderer
        if _handle is None:
            h = C.c_void_p()
            check(lib().spl_scene_create(renderer.handle, int(splat_count), C.byref(h)))
            _handle = h
        self._h = _handle
        c = C.c_uint64()
        check(lib().spl_scene_splat_count(self._h, C.byref(c)))
        self.splat_count = int(c.value)

    @property
    def handle(self):
        return self._h

    @classmethod
    def new_dev_test(cls, renderer: Renderer) -> "Scene":
        h = C.c_void_p()
        check(lib().spl_scene_new_dev_test(renderer.handle, C.byref(h)))
        return cls(renderer, 3, _handle=h)

    @staticmethod
    def parse_file_header(path: str):
        hs, cnt = C.c_uint16(), C.c_uint64()
        check(lib().spl_parse_file_header(path.encode(), C.byref(hs), C.byref(cnt)))
        return int(hs.value), int(cnt.value)

    def load_chunk(self, path: str, file_header_size: int, begin: int, end: int) -> None:
        check(lib().spl_scene_load_chunk(self._h, path.encode(), int(file_header_size), int(begin), int(end)))

    def load_records(self, first: int, records: np.ndarray) -> None:
        rec = np.ascontiguousarray(records)
        if rec.dtype != SERIALIZED_SPLAT_DTYPE:
            rec = rec.view(SERIALIZED_SPLAT_DTYPE) if rec.dtype.itemsize == 248 else rec.astype(SERIALIZED_SPLAT_DTYPE)
        check(lib().spl_scene_load_records(self._h, int(first), int(rec.shape[0]), rec.ctypes.data_as(C.c_void_p)))

    def close(self) -> None:
        if self._h:
            lib().spl_scene_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Serving form (spl_server_*, include/splatter.h): resident scenes, a
    request queue admitted by a device-memory budget, continuous batching of
    same-scene requests over max_in_flight concurrent streams.  A request's
    image equals Renderer.render_frame's for the same scene, matrices and
    size.  config=None takes the library's serving defaults (8 views per
    batch, 8 tile pairs per splat, scenes up to 2^23 splats)."""

    SERVER_DONE, SERVER_PENDING, SERVER_RUNNING = 0, 1, 2

    def __init__(self, device: Device, config: Configuration | None = None, max_in_flight: int = 3,
                 max_width: int = 3840, max_height: int = 2160, memory_budget_bytes: int = 64 << 30):
        c = SplServerConfiguration()
        lib().spl_server_configuration_default(C.byref(c))
        if config is not None:
            c.renderer = config.to_c()
        c.max_in_flight = int(max_in_flight)
        c.max_width, c.max_height = int(max_width), int(max_height)
        c.memory_budget_bytes = int(memory_budget_bytes)
        h = C.c_void_p()
        check(lib().spl_server_create(device.handle, C.byref(c), C.byref(h)))
        self.device = device
        self._h = h

    @property
    def handle(self):
        return self._h

    def create_scene(self, splat_count: int) -> Scene:
        h = C.c_void_p()
        check(lib().spl_server_scene_create(self._h, int(splat_count), C.byref(h)))
        return Scene(self, int(splat_count), _handle=h)

    def submit(self, scene: Scene, width: int, height: int, camera_matrix, view_matrix) -> int:
        t = C.c_uint64()
        check(lib().spl_server_submit(self._h, scene.handle, int(width), int(height), _mat16(camera_matrix),
                                      _mat16(view_matrix), C.byref(t)))
        return int(t.value)

    def poll(self, capacity: int = 4096) -> list:
        buf = (C.c_uint64 * capacity)()
        n = C.c_uint32()
        check(lib().spl_server_poll(self._h, buf, capacity, C.byref(n)))
        return [int(buf[i]) for i in range(n.value)]

    def wait(self, ticket: int) -> None:
        check(lib().spl_server_wait(self._h, int(ticket)))

    def info(self, ticket: int) -> dict:
        i = SplRequestInfo()
        check(lib().spl_server_request_info(self._h, int(ticket), C.byref(i)))
        return {"status": i.status, "width": i.width, "height": i.height, "batch_views": i.batch_views,
                "submit_ms": i.submit_ms, "start_ms": i.start_ms, "done_ms": i.done_ms}

    def read(self, ticket: int) -> np.ndarray:
        info = self.info(ticket)
        out = np.empty((info["height"], info["width"], 4), np.uint8)
        check(lib().spl_server_read(self._h, int(ticket), out.ctypes.data_as(C.c_void_p), out.nbytes))
        return out

    def release(self, ticket: int) -> None:
        check(lib().spl_server_release(self._h, int(ticket)))

    def stats(self) -> dict:
        st = SplServerStats()
        check(lib().spl_server_get_stats(self._h, C.byref(st)))
        return {k: int(getattr(st, k)) for k, _ in SplServerStats._fields_ if k != "reserved"}

    def close(self) -> None:
        if self._h:
            lib().spl_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
