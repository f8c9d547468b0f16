"""GPU parity of the HIP pipeline against the CPU oracle, through the C ABI.

Bars (stated here, checked below):
  visible set + draw order   bit-exact vs the CPU depth-sort fallback
                             (renderer.rs:372-396): keys and splat indices
  contour conic              <= 1e-5 normwise relative error vs the FP64 oracle
  rectangle record           <= 1e-5 normwise relative error vs the FP64 oracle
  image                      <= 2/255 max per-channel error vs the oracle raster
"""
import math

import numpy as np
import pytest

import oracle_bind as ob
import splatter_amd as sa
from gpu_common import Case, config, device, image_diff, normwise_rel, say
from splatter_amd.synth import Workload, load_scene, look_at, orbit_poses, sphere_pose

pytestmark = pytest.mark.gpu

CONIC_TOL = 1e-5
RECORD_TOL = 1e-5
IMAGE_TOL = 2


def check_case(case: Case, conics: bool = True, image: bool = True):
    if conics:
        case.renderer.enable_conics(True)
    img = case.render()
    keys, idx = case.renderer.read_entries()
    okeys, oidx, orec, oimg = case.oracle(with_image=image)
    say(f"  visible gpu={len(idx)} oracle={len(oidx)}")
    assert np.array_equal(idx, oidx)
    if okeys is not None:
        assert np.array_equal(keys, okeys)
    rec = case.renderer.read_records()
    assert rec.shape == orec.shape
    if len(rec):
        assert np.array_equal(rec[:, :4], orec[:, :4])  # colour: same f32 ops
        fin = np.all(np.isfinite(orec[:, 4:]), axis=1)
        assert np.array_equal(fin, np.all(np.isfinite(rec[:, 4:]), axis=1))
        assert normwise_rel(rec[fin, 4:], orec[fin, 4:]).max() <= RECORD_TOL
    if conics and len(idx):
        con = case.renderer.read_conics()
        sub = np.linspace(0, len(idx) - 1, min(len(idx), 2000)).astype(np.int64)
        _, ocon = ob.project(case.splats[oidx[sub]], case.u, case.cfg.spherical_harmonics_order,
                             case.cfg.use_covariance_for_scale)
        assert normwise_rel(con[sub], ocon).max() <= CONIC_TOL
    if image:
        d = image_diff(img, oimg)
        say(f"  image max diff {d}/255, exact fraction {np.mean(img == oimg):.6f}")
        assert d <= IMAGE_TOL
    return img


def test_dev_test_scene():
    r = sa.Renderer(device(), config(max_splat_count=16))
    s = sa.Scene.new_dev_test(r)
    f = sa.Frame(device(), 320, 240)
    cam, view = look_at((0, 0, -5), (0, 0, 0))
    r.render_frame(f, cam, view, s)
    keys, idx = r.read_entries()
    assert list(idx) == [0, 1, 2]
    assert keys[0] == keys[1] < keys[2]
    u = ob.uniforms(cam, view, 320, 240)
    okeys, oidx, orec, oimg = ob.render_frame(ob.dev_test_scene(), u, 3, True)
    assert np.array_equal(keys, okeys)
    img = f.read()[..., [2, 1, 0, 3]]
    assert image_diff(img, oimg) <= IMAGE_TOL


def test_gaussian_100k_720p_fov60():
    """100k Gaussians, camera at distance 5 looking at the origin, 60 deg FOV."""
    case = Case(Workload("gaussian", 100_000, seed=0), config(field_of_view_y=math.radians(60)),
                1280, 720, look_at((0, 0, -5), (0, 0, 0)))
    check_case(case)


def test_gaussian_1m_1080p_random_pose():
    """1M Gaussians, 1920x1080, pose on the radius-6 sphere."""
    case = Case(Workload("gaussian", 1_000_000, seed=1), config(), 1920, 1080, sphere_pose(1))
    check_case(case)


@pytest.mark.parametrize("unaligned,blend", [(True, sa.BlendMode.UNORM8), (False, sa.BlendMode.UNORM8),
                                             (True, sa.BlendMode.F32), (False, sa.BlendMode.F32)])
def test_single_view_split_heavy_tiles(unaligned, blend):
    """One view in mode 3 gives each tile with a list of 8x the mean (and
    2048 pairs) four blocks, one per quadrant: the frame still equals the
    oracle's, and equals the same view rendered in a batch (no split)."""
    say(f"split heavy tiles, unaligned {unaligned}, blend {blend}")
    case = Case(Workload("tight", 200_000, seed=12),
                config(use_unaligned_rectangles=unaligned, blend_mode=blend, raster_mode=3, max_views=2),
                800, 600, orbit_poses(8, radius=8)[5])
    case.renderer.enable_tile_stats(True)
    img = check_case(case, conics=False)
    lens = case.renderer.tile_stats()[:, 0].astype(np.int64)
    heavy = np.count_nonzero(lens >= max(8 * (lens.sum() // len(lens)), 2048))
    say(f"  tiles {len(lens)}, mean list {lens.mean():.0f}, max {lens.max()}, split tiles {heavy}")
    assert heavy > 0
    other = sa.Frame(case.frame.device, case.width, case.height)
    case.renderer.render_views([case.frame, other], [case.cam, orbit_poses(8, radius=8)[1][0]],
                               [case.view, orbit_poses(8, radius=8)[1][1]], case.scene)
    batch = case.frame.read()
    if case.cfg.frame_format == sa.FrameFormat.BGRA8_UNORM:
        batch = batch[..., [2, 1, 0, 3]]
    assert np.array_equal(batch, img)


def test_tie_heavy_tight_clusters():
    """Dense clusters: many equal f32 depths -> ties broken by splat index."""
    case = Case(Workload("tight", 300_000, seed=2), config(), 1280, 720, orbit_poses(8, radius=8)[3])
    img = check_case(case, conics=False)
    keys, _ = case.renderer.read_entries()
    assert np.count_nonzero(np.diff(keys.astype(np.int64)) == 0) > 1000


@pytest.mark.parametrize("sh,cov,unaligned,blend,fmt", [
    (0, True, True, sa.BlendMode.UNORM8, sa.FrameFormat.RGBA8_UNORM),
    (1, False, True, sa.BlendMode.UNORM8, sa.FrameFormat.BGRA8_UNORM),
    (2, True, False, sa.BlendMode.F32, sa.FrameFormat.BGRA8_UNORM),
    (3, False, False, sa.BlendMode.F32, sa.FrameFormat.RGBA8_UNORM),
])
def test_configuration_variants(sh, cov, unaligned, blend, fmt):
    case = Case(Workload("gaussian", 60_000, seed=3),
                config(spherical_harmonics_order=sh, use_covariance_for_scale=cov,
                       use_unaligned_rectangles=unaligned, blend_mode=blend, frame_format=fmt),
                640, 480, sphere_pose(7, radius=5))
    check_case(case)


def test_depth_sorting_none_draws_in_index_order():
    case = Case(Workload("gaussian", 50_000, seed=4), config(depth_sorting=sa.DepthSorting.NONE),
                640, 360, sphere_pose(2))
    check_case(case, conics=False)


@pytest.mark.parametrize("bits,tile,mode,batch,longest", [
    (10, 4096, 0, 256, 0), (8, 8192, 0, 128, 1), (10, 8192, 1, 128, 0), (8, 4096, 1, 256, 1),
    (8, 16384, 2, 128, 0), (10, 16384, 2, 256, 1), (10, 16384, 3, 128, 0), (8, 8192, 3, 256, 1)])
def test_autotuner_options(bits, tile, mode, batch, longest):
    # every option the autotuner may pick gives the same bit-exact order and image
    say(f"radix bits {bits}, sort tile {tile}, raster mode {mode}, batch {batch}, longest first {longest}")
    case = Case(Workload("clustered", 300_000, seed=5),
                config(radix_bits_per_digit=bits, sort_tile_keys=tile, raster_mode=mode, raster_batch=batch,
                       raster_longest_first=longest), 800, 600, sphere_pose(3))
    check_case(case, conics=False)


@pytest.mark.parametrize("mode,unaligned,blend", [(0, True, sa.BlendMode.F32), (0, False, sa.BlendMode.UNORM8),
                                                  (1, True, sa.BlendMode.F32), (1, False, sa.BlendMode.UNORM8),
                                                  (2, True, sa.BlendMode.F32), (2, False, sa.BlendMode.UNORM8),
                                                  (3, True, sa.BlendMode.F32), (3, False, sa.BlendMode.UNORM8)])
def test_raster_modes_other_rectangles_and_blends(mode, unaligned, blend):
    case = Case(Workload("clustered", 300_000, seed=8),
                config(raster_mode=mode, use_unaligned_rectangles=unaligned, blend_mode=blend), 640, 480,
                sphere_pose(8))
    check_case(case, conics=False)


def test_wide_frame_uses_pixel_boxes():
    # more than 255 screen tiles across: binning reads 8-byte pixel boxes
    case = Case(Workload("gaussian", 100_000, seed=9), config(), 4200, 96, sphere_pose(9, radius=3))
    check_case(case, conics=False)


@pytest.mark.parametrize("scale,search", [(1.0, False), (12.0, True)])
def test_bin_owner_table_and_search(scale, search):
    """The binning expands a block's 1024 entries (in depth order) through an
    owner table when they emit at most 4096 pairs, and by binary search over
    their offsets when they emit more.  More pairs than 4096 per block of
    the frame put at least one block above 4096; fewer than four per splat
    put at least one at or below it.  Either way the frame equals the
    oracle's."""
    say(f"splat scale {scale}")
    case = Case(Workload("gaussian", 60_000, seed=21), config(splat_scale=scale), 640, 480,
                sphere_pose(21, radius=4))
    check_case(case, conics=False)
    pairs, visible = case.renderer.pair_count(), case.renderer.visible_count()
    say(f"  pairs {pairs}, visible {visible}, {pairs / max(visible, 1):.2f} a splat")
    assert visible > 4096
    blocks = -(-visible // 1024)
    assert (pairs > 4096 * blocks) if search else (pairs < 4 * visible)


def test_edge_cases_empty_culled_tiny_ragged():
    # empty scene
    case = Case(Workload("gaussian", 0, seed=0), config(max_splat_count=1), 64, 48, sphere_pose(0))
    img = case.render()
    assert case.renderer.visible_count() == 0
    assert (img[..., :3] == 0).all() and (img[..., 3] == 255).all()
    # everything behind the camera
    cam, view = look_at((0, 0, 30), (0, 0, 60))
    case = Case(Workload("gaussian", 5000, seed=1), config(), 64, 48, (cam, view))
    img = case.render()
    assert case.renderer.visible_count() == 0 and (img[..., 3] == 255).all()
    # one splat, and ragged sizes around the 4096-entry tiles and 16-px tiles
    for n, (w, h) in ((1, (17, 9)), (4095, (100, 37)), (4097, (33, 65)), (8191, (255, 1))):
        case = Case(Workload("gaussian", n, seed=n), config(), w, h, sphere_pose(n, radius=4))
        check_case(case, conics=False)


def test_malformed_records():
    """Records a PLY file may hold but no trainer writes: zero, infinite and
    NaN rotations, centres, scales, opacity logits and colours, splats at the
    camera and splats larger than the view.  The decode (scene.rs:225-249)
    turns them into NaN or infinite attributes; the visible set and order
    stay bit-exact, non-finite rectangles get no tiles, and the image stays
    within the bar."""
    w = Workload("gaussian", 20_000, seed=21)
    cam, view = sphere_pose(21, radius=4)
    case = Case(Workload("gaussian", 0, seed=0), config(), 320, 240, (cam, view))
    rec = w.all().copy()
    rec["rotation"][0:100] = 0.0
    rec["center"][100:200] = np.nan
    rec["scale"][200:300] = -np.inf
    rec["scale"][300:400] = 1e3
    rec["alpha"][400:450] = 1e30
    rec["alpha"][450:500] = -1e30
    rec["color"][500:600] = np.nan
    rec["color"][600:650] = 1e30
    rec["scale"][650:660] = 3.0
    rec["alpha"][650:660] = -3.0
    rec["center"][660:700] = cam[3, :3]
    rec["scale"][700:800, 0] = -30.0
    rec["rotation"][800:900] = np.inf
    case.records, case.splats = rec, ob.decode(rec)
    case.scene = sa.Scene(case.renderer, len(rec))
    case.scene.load_records(0, rec)
    img = case.render()
    keys, idx = case.renderer.read_entries()
    okeys, oidx, orec, oimg = case.oracle()
    say(f"  visible gpu={len(idx)} oracle={len(oidx)}")
    assert np.array_equal(idx, oidx) and np.array_equal(keys, okeys)
    got = case.renderer.read_records()
    assert np.array_equal(got[:, :4], orec[:, :4], equal_nan=True)
    fin = np.all(np.isfinite(orec[:, 4:]), axis=1)
    assert (~fin).sum() > 300 and np.isnan(orec[:, :4]).any()
    assert np.array_equal(fin, np.all(np.isfinite(got[:, 4:]), axis=1))
    assert normwise_rel(got[fin, 4:], orec[fin, 4:]).max() <= RECORD_TOL
    d = image_diff(img, oimg)
    say(f"  image max diff {d}/255")
    assert d <= IMAGE_TOL and (img[..., 3] < 255).mean() > 0.2


def test_deterministic_and_frame_size_changes():
    w = Workload("clustered", 400_000, seed=6)
    case = Case(w, config(), 1920, 1080, orbit_poses(16)[5])
    a = case.render()
    ka, ia = case.renderer.read_entries()
    small = sa.Frame(device(), 320, 200)
    big = sa.Frame(device(), 3840, 2160)
    case.renderer.render_frame(big, case.cam, case.view, case.scene)
    case.renderer.render_frame(small, case.cam, case.view, case.scene)
    b = case.render()
    kb, ib = case.renderer.read_entries()
    assert np.array_equal(a, b) and np.array_equal(ka, kb) and np.array_equal(ia, ib)


@pytest.mark.parametrize("graphs,longest,mode", [(True, 0, 3), (False, 0, 3), (True, 1, 3), (True, 0, 0)])
def test_render_views_equals_single_frames(graphs, longest, mode):
    """Batched form: uneven frame sizes in one launch sequence, more views
    than max_views (two batches), each view equal to its single render and
    to the oracle (entries bit-exact, image within tolerance)."""
    w = Workload("clustered", 300_000, seed=8)
    case = Case(w, config(max_views=4, use_hip_graphs=graphs, raster_longest_first=longest, raster_mode=mode),
                640, 360, orbit_poses(6)[0])
    poses = orbit_poses(6)
    sizes = [(640, 360), (320, 240), (1280, 720), (97, 61), (640, 360), (500, 500)]
    frames = [sa.Frame(device(), wd, ht) for wd, ht in sizes]
    for rep in range(2):  # second replay of the same shapes
        case.renderer.render_views(frames, [p[0] for p in poses], [p[1] for p in poses], case.scene)
    batched = [f.read() for f in frames]
    # entries of the views of the second batch (views 4, 5 -> slots 0, 1)
    for slot, vi in ((0, 4), (1, 5)):
        case.renderer.select_view(slot)
        keys, idx = case.renderer.read_entries()
        cam, view = poses[vi]
        u = ob.uniforms(cam, view, *sizes[vi])
        okeys, oidx = ob.cpu_depth_sort(case.splats["center"], ob.vp_of(u), 1.1)
        assert np.array_equal(idx, oidx) and np.array_equal(keys, okeys)
    case.renderer.select_view(0)
    for (cam, view), img, f in zip(poses, batched, frames):
        case.renderer.render_frame(f, cam, view, case.scene)
        assert np.array_equal(f.read(), img)
    cam, view = poses[2]
    u = ob.uniforms(cam, view, *sizes[2])
    _, _, _, oimg = ob.render_frame(case.splats, u, 3, True)
    assert image_diff(batched[2][..., [2, 1, 0, 3]], oimg) <= IMAGE_TOL


def test_batches_in_flight_on_async_renderers():
    """Two renderers on streams of their own, alternating batches with their
    own frames (two batches on the device at once), give exactly the images
    of one renderer on the device stream."""
    w = Workload("clustered", 200_000, seed=10)
    cfg = config(max_views=4)
    poses = orbit_poses(12)
    sizes = [(320, 240), (640, 360), (97, 61), (320, 240)]
    ref_r = sa.Renderer(device(), cfg)
    ref_s = sa.Scene(ref_r, w.count)
    load_scene(ref_s, w)
    ref = []
    for k in range(3):
        fr = [sa.Frame(device(), wd, ht) for wd, ht in sizes]
        sel = poses[4 * k:4 * k + 4]
        ref_r.render_views(fr, [p[0] for p in sel], [p[1] for p in sel], ref_s)
        ref.append([f.read() for f in fr])
    rs = [sa.Renderer(device(), cfg, own_stream=True) for _ in range(2)]
    scene = sa.Scene(rs[0], w.count)
    load_scene(scene, w)
    fsets = [[sa.Frame(device(), wd, ht) for wd, ht in sizes] for _ in range(3)]
    for rep in range(2):  # second round replays the captured graphs
        for k in range(3):
            sel = poses[4 * k:4 * k + 4]
            rs[k % 2].render_views(fsets[k], [p[0] for p in sel], [p[1] for p in sel], scene)
        for r in rs:
            r.synchronize()
        for k in range(3):
            for got, want in zip((f.read() for f in fsets[k]), ref[k]):
                assert np.array_equal(got, want)


def test_batch_wider_than_sixteen_views():
    """max_views = 32: the cull's 32-bit visibility masks and the
    projection's second half of views (views 16..31); every view of a
    20-view batch equals its single render, entries and image."""
    w = Workload("gaussian", 60_000, seed=12)
    poses = orbit_poses(20, radius=4)
    r = sa.Renderer(device(), config(max_views=32, max_splat_count=w.count))
    scene = sa.Scene(r, w.count)
    load_scene(scene, w)
    frames = [sa.Frame(device(), 160 + 8 * (i % 3), 120) for i in range(20)]
    r.render_views(frames, [p[0] for p in poses], [p[1] for p in poses], scene)
    batched = [f.read() for f in frames]
    entries = []
    for vi in (0, 15, 16, 19):
        r.select_view(vi)
        entries.append((vi, *r.read_entries()))
    r.select_view(0)
    for vi, keys, idx in entries:
        cam, view = poses[vi]
        u = ob.uniforms(cam, view, frames[vi].width, frames[vi].height)
        okeys, oidx = ob.cpu_depth_sort(ob.decode(w.all())["center"], ob.vp_of(u), 1.1)
        assert np.array_equal(idx, oidx) and np.array_equal(keys, okeys)
    for (cam, view), img, f in zip(poses, batched, frames):
        r.render_frame(f, cam, view, scene)
        assert np.array_equal(f.read(), img)
