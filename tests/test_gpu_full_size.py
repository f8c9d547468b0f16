"""Parity at the measured workloads' full size, through the C ABI.

- Batched: 6M clustered splats (64 anisotropic clusters + ground slab), SH3,
  16 views of the 256-pose orbit at 1920x1080 in one launch sequence, with
  the options the autotuner picks on MI355X (10-bit digits, 16384-key sort
  tiles, 128-entry raster batches).  Every view's (key, splat index) list is
  bit-exact against the oracle's CPU depth-sort fallback
  (renderer.rs:372-396); two views' images are within 2/255 of the oracle
  raster and equal to their single-view renders byte for byte.
- Serving: 8 resident clustered scenes of 0.3M..6M splats, a burst of
  requests at 640x480..3840x2160 so that same-scene requests batch.  Every
  request is byte-equal mixed, alone and through spl_render_frame; one
  request per resolution has its entries bit-exact against the oracle.
"""
import time

import numpy as np
import pytest

import oracle_bind as ob
import splatter_amd as sa
from gpu_common import config, device, image_diff, say
from splatter_amd.synth import Workload, orbit_poses, sphere_pose

pytestmark = pytest.mark.gpu

IMAGE_TOL = 2


def _load(scene, workload):
    """Uploads the workload; returns its records (decoded by the oracle only
    where a test needs the full splats)."""
    parts = []
    for first, rec in workload.chunks():
        scene.load_records(first, rec)
        parts.append(rec)
    return np.concatenate(parts)


def test_6m_clustered_1080p_16_view_batch():
    t0 = time.time()
    n, w, h, views = 6_000_000, 1920, 1080, 16
    cfg = config(max_splat_count=n, max_views=views, radix_bits_per_digit=10, sort_tile_keys=16384,
                 raster_batch=128)
    r = sa.Renderer(device(), cfg)
    scene = sa.Scene(r, n)
    rec = _load(scene, Workload("clustered", n, seed=2))
    centers = np.ascontiguousarray(rec["center"])
    say(f"6M clustered: loaded in {time.time() - t0:.1f}s")
    orbit = orbit_poses(256)
    poses = [orbit[k] for k in range(0, 256, 16)]  # 16 views spread round the orbit
    frames = [sa.Frame(device(), w, h) for _ in range(views)]
    for rep in range(2):  # capture, then a replay of the captured sequence
        r.render_views(frames, [p[0] for p in poses], [p[1] for p in poses], scene)
    batched = [f.read() for f in frames]
    for v, (cam, view) in enumerate(poses):
        r.select_view(v)
        keys, idx = r.read_entries()
        u = ob.uniforms(cam, view, w, h)
        okeys, oidx = ob.cpu_depth_sort(centers, ob.vp_of(u), cfg.frustum_culling_tolerance)
        say(f"  view {v}: visible gpu={len(idx)} oracle={len(oidx)}")
        assert np.array_equal(idx, oidx) and np.array_equal(keys, okeys), f"view {v}: entries differ"
    r.select_view(0)
    splats = ob.decode(rec)
    del rec
    single = sa.Frame(device(), w, h)
    for v in (0, 11):
        cam, view = poses[v]
        r.render_frame(single, cam, view, scene)
        assert np.array_equal(single.read(), batched[v]), f"view {v}: batched and single-view images differ"
        u = ob.uniforms(cam, view, w, h)
        _, _, _, oimg = ob.render_frame(splats, u, 3, cfg.use_covariance_for_scale)
        d = image_diff(batched[v][..., [2, 1, 0, 3]], oimg)
        say(f"  view {v}: image max diff {d}/255")
        assert d <= IMAGE_TOL
    say(f"6M clustered: done in {time.time() - t0:.1f}s")


SCENE_SIZES = [300_000, 600_000, 1_000_000, 1_500_000, 2_000_000, 3_000_000, 4_500_000, 6_000_000]
RESOLUTIONS = [(640, 480), (1280, 720), (1920, 1080), (3840, 2160)]


def test_serving_8_scenes_up_to_4k():
    t0 = time.time()
    cfg = config(max_splat_count=max(SCENE_SIZES), max_views=8, tile_pairs_per_splat=8)
    budget = 96 << 30
    server = sa.Server(device(), cfg, max_in_flight=3, max_width=3840, max_height=2160,
                       memory_budget_bytes=budget)
    # the plain path's renderer and its copies of the scenes, for the
    # comparison at the end
    r = sa.Renderer(device(), config(max_splat_count=max(SCENE_SIZES), tile_pairs_per_splat=8))
    scenes, plain_scenes, centers = [], [], []
    for i, n in enumerate(SCENE_SIZES):
        sc, pc = server.create_scene(n), sa.Scene(r, n)
        parts = []
        for first, rec in Workload("clustered", n, seed=100 + i).chunks():
            sc.load_records(first, rec)
            pc.load_records(first, rec)
            parts.append(rec["center"])
        centers.append(np.ascontiguousarray(np.concatenate(parts)))
        scenes.append(sc)
        plain_scenes.append(pc)
    device().synchronize()
    say(f"serving: 8 scenes resident in {time.time() - t0:.1f}s")
    rng = np.random.default_rng(21)
    reqs = []
    for i in range(40):
        s = int(rng.integers(len(scenes)))
        reqs.append((s, *RESOLUTIONS[i % len(RESOLUTIONS)], sphere_pose(21_000 + i)))
    # the first pass creates the slots, frames and captured sequences, so the
    # measured burst is submitted faster than the device renders it
    for t in [server.submit(scenes[s], w, h, cam, view) for s, w, h, (cam, view) in reqs]:
        server.wait(t)
        server.release(t)
    warm = server.stats()
    tickets = [server.submit(scenes[s], w, h, cam, view) for s, w, h, (cam, view) in reqs]
    mixed, batch_views = [], []
    for t in tickets:
        server.wait(t)
        info = server.info(t)
        assert info["status"] == sa.Server.SERVER_DONE
        batch_views.append(info["batch_views"])
        mixed.append(server.read(t))
        server.release(t)
    st = server.stats()
    say(f"serving burst: {st['batches'] - warm['batches']} batches for {len(reqs)} requests, "
        f"mean batch {np.mean(batch_views):.2f}, peak {st['memory_peak'] / 2**30:.1f} GiB")
    assert st["failed"] == 0
    assert np.mean(batch_views) > 1.0  # requests that waited went out together
    assert st["memory_peak"] <= budget
    for i, (s, w, h, (cam, view)) in enumerate(reqs):
        t = server.submit(scenes[s], w, h, cam, view)
        alone = server.read(t)
        assert server.info(t)["batch_views"] == 1
        server.release(t)
        assert np.array_equal(mixed[i], alone), f"request {i}: mixed and alone differ"
    server.close()
    # the plain path on the same scenes: byte-equal images, and one request
    # per resolution with its entries against the oracle
    checked = set()
    for i, (s, w, h, (cam, view)) in enumerate(reqs):
        f = sa.Frame(device(), w, h)
        r.render_frame(f, cam, view, plain_scenes[s])
        assert np.array_equal(f.read(), mixed[i]), f"request {i}: server and render_frame differ"
        if (w, h) not in checked:
            checked.add((w, h))
            keys, idx = r.read_entries()
            u = ob.uniforms(cam, view, w, h)
            okeys, oidx = ob.cpu_depth_sort(centers[s], ob.vp_of(u), cfg.frustum_culling_tolerance)
            say(f"  request {i} ({SCENE_SIZES[s]} splats, {w}x{h}): visible gpu={len(idx)} oracle={len(oidx)}")
            assert np.array_equal(idx, oidx) and np.array_equal(keys, okeys)
        f.close()
    assert len(checked) == len(RESOLUTIONS)
    say(f"serving: done in {time.time() - t0:.1f}s")
