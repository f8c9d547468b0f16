"""Pins of the oracle's projection (shaders.wgsl:63-246, 530-547).

- sphere on the optical axis: contour circle radius s/sqrt(d^2-s^2), EWA
  covariance semi-axes s/d (closed forms);
- the contour conic equals the ray-ellipsoid tangency discriminant computed
  directly (independent algebra) for random ellipsoids and poses;
- quaternion convention: a 90 degree turn about z swaps the x/y scales;
- the record's rectangle (translation, axis directions, semi-axes) equals the
  conic's centre and eigen-decomposition from numpy (textbook routines);
- FP32 restatement agrees with FP64 on the covariance path;
- off the optical axis (x/z, y/z beyond the +-1 clamp too), the covariance
  path's semi-axes equal an independent numpy EWA (eigenvalues of
  J C^T R S^2 R^T C J^T), with planted mistakes (sign of J[0][2], no clamp)
  shown to miss.
"""
import math

import numpy as np
import pytest

import oracle_bind as ob
from splatter_amd.synth import Workload, look_at, sphere_pose


def identity_uniforms(w=800, h=800):
    cam = np.eye(4, dtype=np.float32)
    return ob.uniforms(cam, cam, w, h)


def splat(center, scale, rot=(1, 0, 0, 0), alpha=1.0):
    s = np.zeros(1, ob.SPLAT_DTYPE)
    s["center"] = center
    s["scale"] = scale
    s["rotation"] = rot
    s["alpha"] = alpha
    return s


def conic_matrix(c):
    xx, yy, zz, xy, xz, yz = c
    return np.array([[xx, xy, xz], [xy, yy, yz], [xz, yz, zz]])


@pytest.mark.parametrize("s,d", [(0.1, 5.0), (0.5, 3.0), (1.0, 20.0)])
def test_sphere_on_axis_closed_forms(s, d):
    u = identity_uniforms()
    rec, con = ob.project(splat((0, 0, d), (s, s, s)), u, 0, covariance=False)
    M = conic_matrix(con[0])
    assert abs(M[0, 1]) < 1e-12 and abs(M[0, 2]) < 1e-12 and abs(M[1, 2]) < 1e-12
    rho2 = -M[2, 2] / M[0, 0]
    s32 = float(np.float32(s))
    assert abs(math.sqrt(rho2) - s32 / math.sqrt(d * d - s32 * s32)) < 1e-12 * max(1, rho2)
    assert abs(M[1, 1] - M[0, 0]) < 1e-12 * abs(M[0, 0])
    # covariance path: semi-axes s/d and s/(2d) for an ellipsoid turned about
    # the view axis (turning keeps the eigenvalues, and gives M.xy != 0)
    q = (math.cos(0.15), 0.0, 0.0, math.sin(0.15))
    rec_c, _ = ob.project(splat((0, 0, d), (s, s * 0.5, s), q), u, 0, covariance=True)
    a = np.hypot(rec_c[0, 4], rec_c[0, 5]) - u.ellipse_size_bias
    b = np.hypot(rec_c[0, 6], rec_c[0, 7]) - u.ellipse_size_bias
    assert sorted([a, b]) == pytest.approx(sorted([s32 / d, float(np.float32(s * 0.5)) / d]), rel=1e-5)


def test_axis_aligned_contour_collapses_the_rectangle():
    """shaders.wgsl:200-214: with M.xy == 0, sign(M.xy) == 0 gives j = k = 0,
    so the reference draws an axis-aligned ellipse as an empty rectangle
    (zero columns with the covariance scale, 0 * inf = NaN with the contour
    scale: either way nothing is covered)."""
    u = identity_uniforms()
    rec, con = ob.project(splat((0.5, 0, 5.0), (0.2, 0.1, 0.05)), u, 0, covariance=False)
    assert con[0, 3] == 0.0
    assert not np.all(np.isfinite(rec[0, 4:8]))
    rec, con = ob.project(splat((0.5, 0, 5.0), (0.2, 0.1, 0.05)), u, 0, covariance=True)
    assert np.all(rec[0, 4:8] == 0.0)
    assert not ob.raster(rec, u)[..., :3].any()


def tangency_form(center, scale, rot, cam, p):
    """Discriminant of |S^-1 R^T (e + l*C p - c)|^2 = 1 in l, per view-plane point p."""
    r, i, j, k = rot
    R = np.array([[1 - 2 * (j * j + k * k), 2 * (i * j - r * k), 2 * (i * k + r * j)],
                  [2 * (i * j + r * k), 1 - 2 * (i * i + k * k), 2 * (j * k - r * i)],
                  [2 * (i * k - r * j), 2 * (j * k + r * i), 1 - 2 * (i * i + j * j)]])
    Cm = cam.astype(np.float64)
    C3 = Cm[:3, :3].T  # columns = camera axes
    e = Cm[3, :3]
    Sinv = np.diag(1 / np.asarray(scale, np.float64))
    a = (Sinv @ R.T @ (C3 @ p.T)).T
    b = Sinv @ R.T @ (e - np.asarray(center, np.float64))
    return np.einsum("ni,i->n", a, b) ** 2 - np.einsum("ni,ni->n", a, a) * (b @ b - 1)


def test_conic_equals_tangency_discriminant():
    rng = np.random.default_rng(3)
    for t in range(40):
        cam, view = sphere_pose(t, radius=6.0)
        u = ob.uniforms(cam, view, 640, 480)
        center = rng.normal(0, 1.0, 3)
        scale = np.exp(rng.normal(-2, 0.5, 3))
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        sp = splat(center.astype(np.float32), scale.astype(np.float32), q.astype(np.float32))
        _, con = ob.project(sp, u, 0, covariance=False)
        M = conic_matrix(con[0])
        p = np.concatenate([rng.uniform(-1.5, 1.5, (50, 2)), np.ones((50, 1))], 1)
        want = tangency_form(sp["center"][0], sp["scale"][0], sp["rotation"][0].astype(np.float64), cam, p)
        got = np.einsum("ni,ij,nj->n", p, M, p)
        assert np.allclose(got, want, rtol=1e-9, atol=1e-9 * np.abs(want).max())


def test_quaternion_turn_about_z_swaps_scales():
    u = identity_uniforms()
    c = math.sqrt(0.5)
    _, m1 = ob.project(splat((0.3, -0.2, 4.0), (0.3, 0.1, 0.2), (c, 0, 0, c)), u, 0, False)
    _, m2 = ob.project(splat((0.3, -0.2, 4.0), (0.1, 0.3, 0.2)), u, 0, False)
    # normwise: sqrt(0.5) rounds in f32, so the rotation is exact to ~1e-7
    assert np.abs(m1 - m2).max() < 1e-6 * np.abs(m2).max()
    # a 90 degree turn about x instead would not match (planted convention mistake)
    _, m3 = ob.project(splat((0.3, -0.2, 4.0), (0.3, 0.1, 0.2), (c, c, 0, 0)), u, 0, False)
    assert np.abs(m3 - m2).max() > 1e-2 * np.abs(m2).max()


def conic_ellipse(M):
    A = M[:2, :2]
    m = M[:2, 2]
    t = -np.linalg.solve(A, m)
    e = M[2, 2] - m @ np.linalg.solve(A, m)
    lam, vec = np.linalg.eigh(A)
    return t, np.sqrt(-e / lam), vec


def test_record_rectangle_is_the_conic_ellipse():
    w = ob.decode(Workload("gaussian", 3000, seed=4).all())
    cam, view = sphere_pose(8)
    u = ob.uniforms(cam, view, 1920, 1080)
    keys, idx = ob.cpu_depth_sort(w["center"], ob.vp_of(u), 1.1)
    rec, con = ob.project(w[idx[:400]], u, 3, covariance=False)
    bias = u.ellipse_size_bias
    checked = 0
    for r, c in zip(rec, con):
        M = conic_matrix(c)
        t, axes, vec = conic_ellipse(M)
        if not np.all(np.isfinite(axes)):
            continue
        assert np.allclose(r[8:10], t, rtol=1e-5, atol=1e-7)
        cols = [r[4:6].astype(np.float64), r[6:8].astype(np.float64)]
        for col in cols:
            length = np.linalg.norm(col) - bias
            d = col / np.linalg.norm(col)
            k = int(np.argmax(np.abs(vec.T @ d)))
            assert abs(abs(vec[:, k] @ d) - 1) < 1e-6
            assert length == pytest.approx(axes[k], rel=2e-5, abs=1e-8)
        checked += 1
    assert checked > 300


def test_fp32_matches_fp64_on_covariance_path():
    w = ob.decode(Workload("gaussian", 2000, seed=6).all())
    cam, view = sphere_pose(5)
    u = ob.uniforms(cam, view, 1280, 720)
    _, idx = ob.cpu_depth_sort(w["center"], ob.vp_of(u), 1.1)
    r64, _ = ob.project(w[idx[:300]], u, 2, covariance=True)
    r32, _ = ob.project(w[idx[:300]], u, 2, covariance=True, f32=True)
    assert np.array_equal(r64[:, :4], r32[:, :4])  # colour is f32 in both
    scale = np.abs(r64[:, 4:]).max(axis=1, keepdims=True)
    assert np.abs(r64[:, 4:] - r32[:, 4:]).max() < 2e-3 * scale.max()


def ewa_semi_axes(center, scale, rot, cam, clamp=True, j02_sign=-1.0):
    """Independent float64 EWA restatement of the covariance scale path
    (shaders.wgsl:99-118, 240-246): world covariance R S^2 R^T, the view
    position from the camera's own axes and eye (not the view matrix), the
    perspective Jacobian at the position with x/z and y/z clamped to
    [-1, 1], and the square roots of the eigenvalues of the projected 2x2
    covariance from numpy's symmetric eigensolver.  clamp and j02_sign exist
    to plant mistakes."""
    r, i, j, k = np.asarray(rot, np.float64)
    R = np.array([[1 - 2 * (j * j + k * k), 2 * (i * j - r * k), 2 * (i * k + r * j)],
                  [2 * (i * j + r * k), 1 - 2 * (i * i + k * k), 2 * (j * k - r * i)],
                  [2 * (i * k - r * j), 2 * (j * k + r * i), 1 - 2 * (i * i + j * j)]])
    S2 = np.diag(np.asarray(scale, np.float64) ** 2)
    sigma = R @ S2 @ R.T
    C3 = cam[:3, :3].astype(np.float64).T  # columns: camera axes in world space
    e = cam[3, :3].astype(np.float64)
    x, y, z = C3.T @ (np.asarray(center, np.float64) - e)
    if clamp:
        x = np.clip(x / z, -1.0, 1.0) * z
        y = np.clip(y / z, -1.0, 1.0) * z
    J = np.array([[1 / z, 0.0, j02_sign * x / (z * z)], [0.0, 1 / z, -y / (z * z)]])
    cov2 = J @ (C3.T @ sigma @ C3) @ J.T
    lam = np.linalg.eigvalsh(cov2)
    return np.sqrt(np.maximum(lam[::-1], 0.0))  # major, minor


def off_axis_splats(rng, cam, n):
    """Splats in front of the camera at view-space x/z in [-1.8, 1.8] and y/z
    in [-1.5, 1.5] (a third of them beyond the clamp), anisotropic, turned."""
    C3 = cam[:3, :3].astype(np.float64).T
    e = cam[3, :3].astype(np.float64)
    out = []
    for _ in range(n):
        z = rng.uniform(2.0, 10.0)
        v = np.array([rng.uniform(-1.8, 1.8) * z, rng.uniform(-1.5, 1.5) * z, z])
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        out.append(splat((C3 @ v + e).astype(np.float32), np.exp(rng.normal(-3.0, 0.6, 3)).astype(np.float32),
                         q.astype(np.float32)))
    return out


def test_covariance_semi_axes_off_axis_match_independent_ewa():
    """The covariance scale path (the showcase default) pinned off the optical
    axis, where the Jacobian's third column and the +-1 clamp matter: the
    record's semi-axes |col| - bias equal the independent EWA restatement's.
    Planted mistakes in the restatement -- a flipped sign of J[0][2] and a
    dropped clamp -- each disagree with the oracle, so the same mistakes in
    the oracle would fail this test."""
    rng = np.random.default_rng(17)
    worst, flipped, unclamped, beyond = 0.0, 0.0, 0.0, 0
    for t in range(6):
        cam, view = sphere_pose(300 + t, radius=6.0)
        u = ob.uniforms(cam, view, 1920, 1080)
        for sp in off_axis_splats(rng, cam, 40):
            rec, _ = ob.project(sp, u, 0, covariance=True)
            r = rec[0].astype(np.float64)
            got = np.array([np.hypot(r[4], r[5]), np.hypot(r[6], r[7])]) - u.ellipse_size_bias
            args = (sp["center"][0], sp["scale"][0], sp["rotation"][0], cam)
            want = ewa_semi_axes(*args)
            worst = max(worst, float(np.abs(got - want).max() / want[0]))
            flipped = max(flipped, float(np.abs(ewa_semi_axes(*args, j02_sign=1.0) - got).max() / want[0]))
            unclamped = max(unclamped, float(np.abs(ewa_semi_axes(*args, clamp=False) - got).max() / want[0]))
            C3 = cam[:3, :3].astype(np.float64).T
            pv = C3.T @ (sp["center"][0].astype(np.float64) - cam[3, :3])
            beyond += int(abs(pv[0] / pv[2]) > 1.0 or abs(pv[1] / pv[2]) > 1.0)
    assert beyond > 50
    assert worst < 1e-5, worst
    assert flipped > 1e-2 and unclamped > 1e-2, (flipped, unclamped)
