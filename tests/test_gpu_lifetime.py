"""Object lifetimes through the C ABI: a device handle may be closed before
the renderers, scenes, frames and servers created on it, and those may be
destroyed in any order (a garbage collector at interpreter exit picks its
own order).  Each object keeps its device alive."""
import numpy as np
import pytest

import splatter_amd as sa
from gpu_common import config, say
from splatter_amd.synth import Workload, load_scene, sphere_pose

pytestmark = pytest.mark.gpu


def _render(dev, w, close_device_first: bool):
    r = sa.Renderer(dev, config(max_splat_count=w.count))
    s = sa.Scene(r, w.count)
    load_scene(s, w)
    f = sa.Frame(dev, 160, 120)
    if close_device_first:
        dev.close()
    cam, view = sphere_pose(11, radius=4)
    r.render_frame(f, cam, view, s)
    img = f.read()
    return img, (r, s, f)


def test_device_closed_before_its_objects():
    w = Workload("gaussian", 20_000, seed=11)
    ref, objs = _render(sa.Device(0), w, close_device_first=False)
    for o in objs:
        o.close()
    for order in ((0, 1, 2), (2, 0, 1), (1, 2, 0)):
        say(f"destroy order {order}")
        img, objs = _render(sa.Device(0), w, close_device_first=True)
        assert np.array_equal(img, ref)
        for i in order:
            objs[i].close()


def test_server_outlives_its_device_handle():
    dev = sa.Device(0)
    cfg = sa.Configuration(spherical_harmonics_order=3, max_splat_count=20_000, max_views=2)
    server = sa.Server(dev, cfg, max_in_flight=2, max_width=320, max_height=240, memory_budget_bytes=4 << 30)
    scene = server.create_scene(20_000)
    load_scene(scene, Workload("gaussian", 20_000, seed=12))
    dev.close()
    cam, view = sphere_pose(12, radius=4)
    req = server.submit(scene, 160, 120, cam, view)
    server.wait(req)
    img = server.read(req)
    assert img.shape == (120, 160, 4) and (img[..., 3] < 255).any()
    server.release(req)
    server.close()
    scene.close()


def test_pair_capacity_overflow_is_reported_and_recoverable():
    """More (splat, tile) pairs than tile_pairs_per_splat * max_splat_count:
    the bin keeps the pairs that fit, the frame completes, pair_count
    reports SPL_ERROR_CAPACITY, and the next frame that fits is exact."""
    dev = sa.Device(0)
    cfg = config(max_splat_count=20_000, tile_pairs_per_splat=1)
    r = sa.Renderer(dev, cfg)
    big = Workload("gaussian", 20_000, seed=31).all().copy()
    big["scale"] += 2.5  # log-scales: every splat spans many 16-pixel tiles
    s_big = sa.Scene(r, len(big))
    s_big.load_records(0, big)
    f = sa.Frame(dev, 1280, 720)
    cam, view = sphere_pose(31, radius=4)
    say("overflowing frame")
    r.render_frame(f, cam, view, s_big)
    with pytest.raises(sa.SplatterError):
        r.pair_count()
    assert f.read().shape == (720, 1280, 4)

    small = Workload("gaussian", 2_000, seed=32)
    s_small = sa.Scene(r, small.count)
    load_scene(s_small, small)
    r.render_frame(f, cam, view, s_small)
    assert 0 < r.pair_count() <= 20_000
    after = f.read()
    fresh = sa.Renderer(dev, cfg)
    fresh.render_frame(f, cam, view, s_small)
    assert np.array_equal(after, f.read())
    for o in (fresh, s_small, s_big, f, r):
        o.close()


def test_destroyed_scene_does_not_replay_its_launch_sequence():
    """A renderer caches one captured launch sequence per batch shape, with the
    scene's device buffers baked in.  Scene A is rendered, destroyed, and
    scene B of the same splat count (different contents) is created on the
    same renderer, often at A's old host address: B's frame must equal a
    fresh renderer's frame of B, several times over."""
    dev = sa.Device(0)
    cfg = config(max_splat_count=30_000)
    r = sa.Renderer(dev, cfg)
    f = sa.Frame(dev, 320, 240)
    cam, view = sphere_pose(41, radius=4)
    wa, wb = Workload("gaussian", 30_000, seed=41), Workload("clustered", 30_000, seed=42)
    for rnd in range(3):
        say(f"destroy-then-recreate round {rnd}")
        a = sa.Scene(r, wa.count)
        load_scene(a, wa)
        r.render_frame(f, cam, view, a)
        img_a = f.read()
        a.close()
        b = sa.Scene(r, wb.count)
        load_scene(b, wb)
        r.render_frame(f, cam, view, b)
        img_b = f.read()
        fresh = sa.Renderer(dev, cfg)
        fresh.render_frame(f, cam, view, b)
        want = f.read()
        fresh.close()
        b.close()
        assert np.array_equal(img_b, want)
        assert not np.array_equal(img_a, want)
    r.close()
    f.close()
