"""spl_render_frame_static (the entry point that allocates nothing) through
the C ABI against spl_render_frame on the same scenes and poses: byte-equal
images on two calls in a row, the second with a smaller scene and frame
(the shared scratch keeps the first call's contents), then a third call
back at the first size."""
import math

import numpy as np
import pytest

import splatter_amd as sa
from gpu_common import config, device, say
from splatter_amd.synth import Workload, load_scene, look_at, sphere_pose

pytestmark = pytest.mark.gpu


def _scene(r, w):
    s = sa.Scene(r, w.count)
    load_scene(s, w)
    return s


def test_static_frame_equals_render_frame():
    dev = device()
    cfg = config(max_splat_count=131_072, field_of_view_y=math.radians(60))
    r = sa.Renderer(dev, cfg)
    cases = [
        (Workload("clustered", 131_072, seed=51), 1920, 1080, sphere_pose(51, radius=5.0)),
        (Workload("gaussian", 20_000, seed=52), 640, 360, look_at((0, 0, -5), (0, 0, 0))),
        (Workload("gaussian", 100_000, seed=53), 1280, 720, sphere_pose(53, radius=6.0)),
    ]
    for i, (w, width, height, (cam, view)) in enumerate(cases):
        say(f"static frame {i}: {w.count} splats at {width}x{height}")
        scene = _scene(r, w)
        f_static, f_plain = sa.Frame(dev, width, height), sa.Frame(dev, width, height)
        dev.render_frame_static(cfg, f_static, cam, view, scene)
        got = f_static.read()
        r.render_frame(f_plain, cam, view, scene)
        want = f_plain.read()
        assert (want[..., 3] < 255).mean() > 0.05
        assert np.array_equal(got, want), f"call {i}: max diff {np.abs(got.astype(int) - want.astype(int)).max()}"
        for o in (f_static, f_plain, scene):
            o.close()
    r.close()


def test_static_frame_rejects_what_exceeds_the_scratch():
    dev = device()
    cfg = config(max_splat_count=200_000)
    r = sa.Renderer(dev, cfg)
    big = sa.Scene(r, 131_073)
    f = sa.Frame(dev, 64, 64)
    cam, view = sphere_pose(1)
    with pytest.raises(sa.SplatterError):
        dev.render_frame_static(cfg, f, cam, view, big)
    small = sa.Scene(r, 10)
    wide = sa.Frame(dev, 3841, 16)
    with pytest.raises(sa.SplatterError):
        dev.render_frame_static(cfg, wide, cam, view, small)
    for o in (wide, small, f, big, r):
        o.close()
