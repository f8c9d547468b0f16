"""Pins of the oracle's raster/fragment/blend (shaders.wgsl:548-592,
renderer.rs:260-283,442).

- an isolated splat's footprint equals the analytic Gaussian truncated at the
  margin rectangle, evaluated in float64;
- two coincident splats follow the closed-form "under" compositing
  C = c1 a1 + (1 - a1) c2 a2, T = (1 - a1)(1 - a2);
- the fragment threshold alpha < 1/255 drops a splat;
- front-to-back: an opaque first splat hides the second;
- the pixel-box walk equals a brute-force walk over every pixel;
- the axis-aligned square (use_unaligned_rectangles = false) equals its
  analytic footprint, with the planted min-for-max radius shown to miss.
"""
import math

import numpy as np

import oracle_bind as ob

W, H = 96, 64


def uni(fov=math.pi / 2):
    cam = np.eye(4, dtype=np.float32)
    return ob.uniforms(cam, cam, W, H, fov=fov)


def record(rgb, alpha, a, b, t, angle=0.0):
    k, j = math.cos(angle), math.sin(angle)
    return np.array([*rgb, alpha, k * a, -j * a, j * b, k * b, *t], np.float32)


def pixel_q(u):
    px = (np.arange(W, dtype=np.float64) + 0.5) / W * 2 - 1
    py = 1 - (np.arange(H, dtype=np.float64) + 0.5) / H * 2
    return px[None, :] * u.view_size[0], py[:, None] * u.view_size[1]


def test_isolated_splat_analytic_footprint():
    u = uni()
    a, b, ang = 0.25, 0.1, 0.4
    rec = record((0.8, 0.4, 0.2), 0.9, a, b, (0.1, -0.05), ang)
    img = ob.raster(rec[None], u, blend_f32=True)
    qx, qy = pixel_q(u)
    dx, dy = qx - 0.1, qy + 0.05
    k, j = math.cos(ang), math.sin(ang)
    tcx = (k * dx - j * dy) / a
    tcy = (j * dx + k * dy) / b
    inside = (np.abs(tcx) < 2.0) & (np.abs(tcy) < 2.0)
    alpha = 0.9 * np.exp(-0.5 * (tcx ** 2 + tcy ** 2))
    drawn = inside & (alpha >= 1 / 255)
    want_r = np.where(drawn, np.round(np.clip(0.8 * alpha, 0, 1) * 255), 0)
    want_t = np.where(drawn, np.round((1 - alpha) * 255), 255)
    assert drawn.sum() > 300
    assert np.abs(img[..., 0].astype(int) - want_r).max() <= 1
    assert np.abs(img[..., 3].astype(int) - want_t).max() <= 1
    # unorm8 emulation: one blend from the clear state rounds the same way
    img8 = ob.raster(rec[None], u, blend_f32=False)
    assert np.abs(img8.astype(int) - img.astype(int)).max() <= 1


def test_two_coincident_splats_closed_form():
    u = uni()
    r1 = record((0.9, 0.1, 0.3), 0.6, 0.3, 0.2, (0.0, 0.0), 0.2)
    r2 = record((0.2, 0.7, 0.5), 0.8, 0.3, 0.2, (0.0, 0.0), 0.2)
    img = ob.raster(np.stack([r1, r2]), u, blend_f32=True)
    qx, qy = pixel_q(u)
    k, j = math.cos(0.2), math.sin(0.2)
    tcx = (k * qx - j * qy) / 0.3
    tcy = (j * qx + k * qy) / 0.2
    g = np.exp(-0.5 * (tcx ** 2 + tcy ** 2))
    inside = (np.abs(tcx) < 2) & (np.abs(tcy) < 2)
    a1 = np.where(inside & (0.6 * g >= 1 / 255), 0.6 * g, 0)
    a2 = np.where(inside & (0.8 * g >= 1 / 255), 0.8 * g, 0)
    for ch, (c1, c2) in enumerate(((0.9, 0.2), (0.1, 0.7), (0.3, 0.5))):
        want = np.round(np.clip(c1 * a1 + (1 - a1) * c2 * a2, 0, 1) * 255)
        assert np.abs(img[..., ch].astype(int) - want).max() <= 1
    want_t = np.round((1 - a1) * (1 - a2) * 255)
    assert np.abs(img[..., 3].astype(int) - want_t).max() <= 1


def test_alpha_threshold_and_front_to_back_occlusion():
    u = uni()
    faint = record((1, 1, 1), 1.0 / 255.0 * 0.99, 0.3, 0.3, (0, 0), 0.3)
    assert not ob.raster(faint[None], u)[..., :3].any()
    opaque = record((1, 0, 0), 1.0, 0.3, 0.3, (0, 0), 0.3)
    green = record((0, 1, 0), 1.0, 0.3, 0.3, (0, 0), 0.3)
    img = ob.raster(np.stack([opaque, green]), u)
    centre = img[H // 2, W // 2].astype(int)
    assert centre[0] >= 250 and centre[3] <= 2 and centre[1] <= 2
    swapped = ob.raster(np.stack([green, opaque]), u)[H // 2, W // 2].astype(int)
    assert swapped[1] >= 250 and swapped[0] <= 2


def test_bbox_walk_equals_brute_force():
    rng = np.random.default_rng(0)
    u = uni()
    recs = []
    for _ in range(300):
        recs.append(record(rng.uniform(-0.2, 1.2, 3), rng.uniform(0, 1), rng.uniform(0.001, 0.3),
                           rng.uniform(0.001, 0.3), rng.uniform(-2, 2, 2), rng.uniform(0, 3)))
    recs = np.stack(recs)
    for unaligned in (True, False):
        for f32 in (False, True):
            a = ob.raster(recs, u, unaligned, f32, brute_force=False)
            b = ob.raster(recs, u, unaligned, f32, brute_force=True)
            assert np.array_equal(a, b)
    assert ob.raster(recs, u)[..., :3].any()


def test_clear_is_black_with_unit_alpha():
    img = ob.raster(np.zeros((0, 10), np.float32), uni())
    assert (img[..., :3] == 0).all() and (img[..., 3] == 255).all()


def square_footprint(u, rec, margin=2.0, radius_of=max):
    """Float64 analytic footprint of an axis-aligned square splat
    (shaders.wgsl:562-571): pixel centres q with |q - t|_inf < radius *
    margin, radius = max(|col0|, |col1|), texture coordinate A^-1 (q - t),
    alpha = a exp(-|tc|^2 / 2).  radius_of plants the min-for-max mistake."""
    a, col0, col1, t = rec[3], rec[4:6].astype(np.float64), rec[6:8].astype(np.float64), rec[8:10]
    qx, qy = pixel_q(u)
    dx, dy = qx - t[0], qy - t[1]
    radius = radius_of(np.linalg.norm(col0), np.linalg.norm(col1))
    inside = (np.abs(dx) < radius * margin) & (np.abs(dy) < radius * margin)
    A = np.stack([col0, col1], axis=1)
    Ai = np.linalg.inv(A)
    tcx = Ai[0, 0] * dx + Ai[0, 1] * dy
    tcy = Ai[1, 0] * dx + Ai[1, 1] * dy
    alpha = a * np.exp(-0.5 * (tcx ** 2 + tcy ** 2))
    return inside & (alpha >= 1 / 255), alpha


def test_axis_aligned_square_analytic_footprint():
    """use_unaligned_rectangles = false: the square quad around an elongated,
    turned splat.  The oracle's image equals the analytic footprint within one
    code; the planted min-for-max radius would crop the splat's long axis and
    miss the oracle by far more."""
    u = uni()
    rec = record((0.8, 0.4, 0.2), 0.9, 0.3, 0.08, (0.1, -0.05), 0.6)
    img = ob.raster(rec[None], u, unaligned=False, blend_f32=True)
    drawn, alpha = square_footprint(u, rec)
    want_r = np.where(drawn, np.round(np.clip(0.8 * alpha, 0, 1) * 255), 0)
    want_t = np.where(drawn, np.round((1 - alpha) * 255), 255)
    assert drawn.sum() > 300
    assert np.abs(img[..., 0].astype(int) - want_r).max() <= 1
    assert np.abs(img[..., 3].astype(int) - want_t).max() <= 1
    # the square covers more than the rotated rectangle would
    assert (img[..., 3] < 255).sum() > (ob.raster(rec[None], u, unaligned=True, blend_f32=True)[..., 3] < 255).sum()
    planted, _ = square_footprint(u, rec, radius_of=min)
    want_planted = np.where(planted, np.round(np.clip(0.8 * alpha, 0, 1) * 255), 0)
    assert np.abs(img[..., 0].astype(int) - want_planted).max() > 20
    img8 = ob.raster(rec[None], u, unaligned=False, blend_f32=False)
    assert np.abs(img8.astype(int) - img.astype(int)).max() <= 1
