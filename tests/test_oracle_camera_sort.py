"""Pins of the oracle's camera math and CPU depth-sort fallback.

renderer.rs:364-396, utils.rs:45-70.  Checks: closed-form projection, the
dev-test scene's known answer (ties broken by index), brute force in float64
away from the culling boundary, invariants of the sorted entries, numpy's
stable argsort as an independent stable sort, and a planted mistake
(division in place of reciprocal-multiply) that the key must not follow.
The vertex-stage predicate of DepthSorting::None (shaders.wgsl:54-61) is
pinned on points at the frustum's edges against an f32 restatement, with the
planted division by h.w alone shown to give a different visible set.
"""
import math

import numpy as np
import pytest

import oracle_bind as ob
from splatter_amd.synth import Workload, look_at, sphere_pose

NEAR, FAR = 1.0, 1000.0


def test_perspective_closed_form():
    out = np.zeros(16, np.float32)
    ob.lib().orc_perspective_projection(1.0, 1.0, 1.0, 1000.0, ob.ptr(out))
    P = out.reshape(4, 4)  # P[c] = column c
    assert P[0, 0] == 1 and P[1, 1] == 1 and P[2, 3] == 1
    assert P[2, 2] == np.float32(1000.0 / 999.0)
    assert abs(P[3, 2] - (-1000.0 / 999.0)) < 1e-6
    assert np.count_nonzero(P) == 5


def test_fov_90_is_the_reference_value_and_60_scales_width():
    cam, view = look_at((0, 0, -5), (0, 0, 0))
    u = ob.uniforms(cam, view, 1920, 1080, fov=math.pi / 2)
    assert u.view_size[1] == 1.0  # tanf(pi/4) == 1.0f: bit-identical to renderer.rs:366-368
    assert u.view_size[0] == np.float32(1920.0 / 1080.0)
    u60 = ob.uniforms(cam, view, 1280, 720, fov=math.radians(60))
    assert abs(u60.view_size[1] - math.tan(math.radians(30))) < 1e-6
    assert abs(u60.view_size[0] - (1280 / 720) * math.tan(math.radians(30))) < 1e-6
    assert u60.ellipse_size_bias == np.float32(np.float32(0.2) * np.float32(u60.view_size[0])) / np.float32(1280)


def test_camera_times_view_is_identity():
    for seed in range(5):
        cam, view = sphere_pose(seed)
        C = cam.astype(np.float64).T  # row-major matrices
        V = view.astype(np.float64).T
        assert np.allclose(C @ V, np.eye(4), atol=1e-5)


def dev_test_keys():
    scene = ob.dev_test_scene()
    cam, view = look_at((0, 0, -5), (0, 0, 0))
    u = ob.uniforms(cam, view, 800, 600)
    keys, idx = ob.cpu_depth_sort(scene["center"], ob.vp_of(u), 1.1)
    return keys, idx


def test_dev_test_scene_known_answer():
    keys, idx = dev_test_keys()
    # all three visible; e_x and e_y share z_view = 5 -> bit-identical keys,
    # so the stable order keeps 0 before 1; e_z is farther (z_view = 6)
    assert list(idx) == [0, 1, 2]
    assert keys[0] == keys[1] and keys[2] > keys[0]
    for k, zv in ((keys[0], 5.0), (keys[2], 6.0)):
        z_ndc = FAR / (FAR - NEAR) - NEAR * FAR / ((FAR - NEAR) * zv)
        assert abs(np.uint32(k).view(np.float32) - z_ndc) < 4e-7


def keys_f64(pos, vp):
    V = vp.astype(np.float64).reshape(4, 4)  # V[c] = column c
    h = pos.astype(np.float64) @ V[:3] + V[3]
    return h[:, :3] / h[:, 3:4], h[:, 3]


def test_brute_force_float64_matches_away_from_boundary():
    rng = np.random.default_rng(1)
    pos = rng.normal(0, 2, size=(4000, 3)).astype(np.float32)
    cam, view = sphere_pose(3)
    u = ob.uniforms(cam, view, 640, 480)
    vp = ob.vp_of(u)
    keys, idx = ob.cpu_depth_sort(pos, vp, 1.1)
    clip, w = keys_f64(pos, vp)
    margin = 1e-4
    vis = (np.abs(clip[:, 0]) < 1.1) & (np.abs(clip[:, 1]) < 1.1) & (np.abs(clip[:, 2] - 0.5) < 0.5)
    clear = ((np.abs(np.abs(clip[:, 0]) - 1.1) > margin) & (np.abs(np.abs(clip[:, 1]) - 1.1) > margin)
             & (np.abs(np.abs(clip[:, 2] - 0.5) - 0.5) > margin))
    got = np.zeros(len(pos), bool)
    got[idx] = True
    assert np.array_equal(got[clear], vis[clear])
    # order = ascending f64 depth (distinct depths)
    z = clip[idx, 2]
    assert np.all(np.diff(z) >= -1e-6)
    assert np.allclose(keys.view(np.float32), clip[idx, 2], rtol=1e-5, atol=1e-6)


def test_invariants_and_stability_on_tie_heavy_input():
    rng = np.random.default_rng(7)
    base = rng.normal(0, 2, size=(500, 3)).astype(np.float32)
    pos = base[rng.integers(0, 500, size=20000)]  # every centre repeated ~40 times
    cam, view = sphere_pose(11)
    vp = ob.vp_of(ob.uniforms(cam, view, 1920, 1080))
    keys, idx = ob.cpu_depth_sort(pos, vp, 1.1)
    assert len(idx) > 1000
    assert np.all(np.diff(keys.astype(np.int64)) >= 0)
    ties = np.diff(keys.astype(np.int64)) == 0
    assert ties.sum() > 1000
    assert np.all(np.diff(idx.astype(np.int64))[ties] > 0)  # stable: index order inside ties
    assert len(np.unique(idx)) == len(idx)
    # the visible set is the filter of renderer.rs:383-385 on the same keys
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(order, np.arange(len(keys)))


def test_merge_sort_equals_numpy_stable_argsort():
    rng = np.random.default_rng(2)
    for n in (0, 1, 2, 3, 17, 1000, 4097):
        k = rng.integers(0, 50, size=n).astype(np.uint32)
        v = np.arange(n, dtype=np.uint32) * 7
        kk, vv = k.copy(), v.copy()
        scratch = np.empty(max(2 * n, 1), np.uint32)
        ob.lib().orc_stable_sort_pairs(ob.ptr(kk), ob.ptr(vv), n, ob.ptr(scratch))
        o = np.argsort(k, kind="stable")
        assert np.array_equal(kk, k[o]) and np.array_equal(vv, v[o])


def test_key_uses_reciprocal_multiply_not_division():
    """Planted mistake: clip = h / h.w instead of h * (1/h.w) (renderer.rs:382).
    The two differ in the last bit for some splats; the oracle must follow
    the reference's spelling."""
    w = Workload("gaussian", 200000, seed=5).all()
    pos = w["center"].astype(np.float32)
    cam, view = sphere_pose(2)
    vp = ob.vp_of(ob.uniforms(cam, view, 1920, 1080))
    keys, idx = ob.cpu_depth_sort(pos, vp, 1.1)
    V = vp.reshape(4, 4)
    p = pos[idx]
    h = [((V[0, r] * p[:, 0] + V[1, r] * p[:, 1]) + V[2, r] * p[:, 2]) + V[3, r] * np.float32(1) for r in range(4)]
    recip = (np.float32(1) / h[3]).astype(np.float32)
    z_mul = (h[2] * recip).astype(np.float32)
    z_div = (h[2] / h[3]).astype(np.float32)
    assert np.array_equal(keys, z_mul.view(np.uint32))
    assert np.count_nonzero(z_mul != z_div) > 0


def test_vertex_stage_predicate_differs_from_fallback_only_at_the_boundary():
    """DepthSorting::None culls with h/(h.w + 1e-7) (shaders.wgsl:54-61)."""
    w = Workload("gaussian", 50000, seed=9).all()
    cam, view = sphere_pose(4)
    vp = ob.vp_of(ob.uniforms(cam, view, 1280, 720))
    a = ob.gpu_predicate_cull(w["center"], vp, 1.1)
    _, b = ob.cpu_depth_sort(w["center"], vp, 1.1)
    assert np.all(np.diff(a.astype(np.int64)) > 0)
    sym = np.setxor1d(a, np.sort(b))
    assert len(sym) <= 3


def test_vertex_stage_predicate_divides_by_w_plus_1e7():
    """DepthSorting::None culls with h.xyz / (h.w + 1e-7) (shaders.wgsl:54-61).
    Points placed on the x and y edges of the frustum at near depths, where
    the 1e-7 moves the quotient by a fraction of an ulp: the oracle's visible
    set is exactly the f32 restatement with the 1e-7, and the planted
    division by h.w alone gives a different set."""
    cam, view = sphere_pose(6, radius=3.0)
    u = ob.uniforms(cam, view, 1280, 720)
    vp = ob.vp_of(u)
    V = vp.reshape(4, 4).astype(np.float64)
    rng = np.random.default_rng(23)
    n = 400_000
    tol = 1.1
    # view-space points on the clip-space edges |c.x| = tol or |c.y| = tol
    cam64 = cam.astype(np.float64)
    C3, e = cam64[:3, :3].T, cam64[3, :3]
    z = rng.uniform(1.05, 1.6, n)
    edge = rng.choice([-1.0, 1.0], n) * tol * (1.0 + rng.normal(0, 3e-7, n))
    other = rng.uniform(-0.9, 0.9, n) * tol
    on_x = rng.uniform(size=n) < 0.5
    cx = np.where(on_x, edge, other)
    cy = np.where(on_x, other, edge)
    vx, vy = cx * z * u.view_size[0], cy * z * u.view_size[1]
    pos = (np.stack([vx, vy, z], 1) @ C3.T + e).astype(np.float32)
    got = ob.gpu_predicate_cull(pos, vp, tol)
    Vf = V.astype(np.float32)
    h = [((Vf[0, r] * pos[:, 0] + Vf[1, r] * pos[:, 1]) + Vf[2, r] * pos[:, 2]) + Vf[3, r] * np.float32(1)
         for r in range(4)]

    def visible(w):
        c = [(h[r] / w).astype(np.float32) for r in range(3)]
        return np.nonzero((np.abs(c[0]) < np.float32(tol)) & (np.abs(c[1]) < np.float32(tol))
                          & (np.abs(c[2] - np.float32(0.5)) < np.float32(0.5)))[0].astype(np.uint32)

    want = visible((h[3] + np.float32(1e-7)).astype(np.float32))
    planted = visible(h[3])
    assert np.array_equal(got, want)
    assert len(np.setxor1d(got, planted)) >= 10, len(np.setxor1d(got, planted))


@pytest.mark.parametrize("threads", [1, 2, 3, 8, 16])
def test_all_core_depth_sort_equals_one_core(threads):
    """The all-cores CPU baseline leg (parallel transform, parallel stable
    merge sort) returns exactly the one-core fallback's entries, on
    tie-heavy input and on the degenerate sizes."""
    w = Workload("tight", 300_000, seed=3).all()
    cam, view = sphere_pose(7, radius=8.0)
    vp = ob.vp_of(ob.uniforms(cam, view, 1920, 1080))
    for n in (0, 1, 5, 4097, 300_000):
        pos = w["center"][:n]
        k1, i1 = ob.cpu_depth_sort(pos, vp, 1.1)
        kt, it = ob.cpu_depth_sort_threads(pos, vp, 1.1, threads)
        assert np.array_equal(k1, kt) and np.array_equal(i1, it), n
    assert len(i1) > 1000 and len(np.unique(k1)) < len(k1)  # ties present
