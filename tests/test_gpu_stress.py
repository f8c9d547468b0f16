"""Sort/raster stress at full size: 32M splats in 16 tight clusters
(sigma 0.05, dense tie-heavy f32 depths), 3840x2160, 8 views batched.
Every view's visible set and stable order are checked bit for bit against
the oracle's CPU depth-sort fallback; the batched images must equal the
single-view path's byte for byte."""
import time

import numpy as np
import pytest

import oracle_bind as ob
import splatter_amd as sa
from gpu_common import config, device, say
from splatter_amd.synth import Workload, orbit_poses

pytestmark = pytest.mark.gpu

N, W, H, VIEWS = 32_000_000, 3840, 2160, 8


def test_32m_tight_clusters_4k_8_views():
    t0 = time.time()
    # dense clusters at 4K: up to ~4 screen tiles per splat
    cfg = config(max_splat_count=N, max_views=VIEWS, sort_tile_keys=8192, tile_pairs_per_splat=8)
    r = sa.Renderer(device(), cfg)
    scene = sa.Scene(r, N)
    centers = np.empty((N, 3), np.float32)
    for first, rec in Workload("tight", N, seed=4).chunks():
        scene.load_records(first, rec)
        centers[first:first + len(rec)] = rec["center"]
    say(f"stress: {N} splats loaded in {time.time() - t0:.1f}s")
    poses = orbit_poses(VIEWS, radius=9.0)
    frames = [sa.Frame(device(), W, H) for _ in range(VIEWS)]
    r.render_views(frames, [p[0] for p in poses], [p[1] for p in poses], scene)
    batched = [f.read() for f in frames]
    ties_total = 0
    for v, (cam, view) in enumerate(poses):
        r.select_view(v)
        keys, idx = r.read_entries()
        u = ob.uniforms(cam, view, W, H)
        okeys, oidx = ob.cpu_depth_sort(centers, ob.vp_of(u), cfg.frustum_culling_tolerance)
        ties = int(len(okeys) - len(np.unique(okeys)))
        ties_total += ties
        say(f"  view {v}: visible gpu={len(idx)} oracle={len(oidx)}, tied keys {ties}, pairs {r.pair_count()}")
        assert np.array_equal(idx, oidx) and np.array_equal(keys, okeys)
    assert ties_total > 0  # the workload really is tie-heavy
    single = sa.Frame(device(), W, H)
    for v in (0, VIEWS - 1):
        r.render_frame(single, poses[v][0], poses[v][1], scene)
        assert np.array_equal(single.read(), batched[v]), f"view {v}: batched and single-view images differ"
    assert batched[0][..., :3].any()
    say(f"stress: done in {time.time() - t0:.1f}s")
