"""Pins of the oracle's record decode (scene.rs:225-249), spherical harmonics
(shaders.wgsl:248-294) and exp2 (the fragment's Gaussian, shaders.wgsl:586)."""
import math

import numpy as np

import oracle_bind as ob
from splatter_amd import SERIALIZED_SPLAT_DTYPE
from splatter_amd.synth import Workload


def test_decode_closed_forms():
    rec = Workload("gaussian", 3000, seed=1).all()
    rec["scale"][0] = (-4.0, -4.0, 20.0)  # outlier: clamped to average + 5
    out = ob.decode(rec)
    q = out["rotation"].astype(np.float64)
    assert np.allclose(np.linalg.norm(q, axis=1), 1.0, atol=3e-7)
    raw = rec["rotation"].astype(np.float64)
    assert np.allclose(q, raw / np.linalg.norm(raw, axis=1, keepdims=True), atol=1e-6)
    s = rec["scale"].astype(np.float64)
    avg = s.mean(axis=1, keepdims=True)
    want = np.exp(np.clip(s, avg - 5, avg + 5))
    assert np.allclose(out["scale"], want, rtol=2e-6)
    assert abs(out["scale"][0, 2] - math.exp((-4 - 4 + 20) / 3 + 5)) < 1e-3
    assert np.allclose(out["alpha"], 1 / (1 + np.exp(-rec["alpha"].astype(np.float64))), rtol=1e-6)
    assert np.array_equal(out["sh"], rec["color"])
    assert np.array_equal(out["center"], rec["center"])


def test_dev_test_scene_values():
    s = ob.dev_test_scene()  # scene.rs:157-192
    assert np.array_equal(s["rotation"], np.tile([1, 0, 0, 0], (3, 1)))
    assert np.array_equal(s["center"], np.eye(3, dtype=np.float32))
    assert np.allclose(s["scale"], np.tile([0.2, 0.1, 0.05], (3, 1)))
    assert np.array_equal(s["alpha"], np.ones(3, np.float32))
    assert np.array_equal(s["sh"][:, :3], np.eye(3, dtype=np.float32))
    assert not s["sh"][:, 3:].any()


def test_sh_constants_closed_forms():
    sp = math.sqrt(math.pi)
    want = [1 / (2 * sp), -math.sqrt(3) / (2 * sp), math.sqrt(3) / (2 * sp), -math.sqrt(3) / (2 * sp),
            math.sqrt(15) / (2 * sp), -math.sqrt(15) / (2 * sp), math.sqrt(5) / (4 * sp),
            -math.sqrt(15) / (2 * sp), math.sqrt(15) / (4 * sp),
            -math.sqrt(35 / 2) / (4 * sp), math.sqrt(105) / (2 * sp), -math.sqrt(21 / 2) / (4 * sp),
            math.sqrt(7) / (4 * sp), -math.sqrt(21 / 2) / (4 * sp), math.sqrt(105) / (4 * sp),
            -math.sqrt(35 / 2) / (4 * sp)]
    for k, w in enumerate(want):
        assert abs(ob.lib().orc_sh_constant(k) - w) < 1e-7 * max(1, abs(w)), k


def sh_basis(d: np.ndarray, order: int) -> np.ndarray:
    """Y_k(d) as the oracle evaluates it: colour - 0.5 with coefficient k = 1."""
    nb = (order + 1) ** 2
    out = np.zeros((len(d), nb))
    col = np.zeros(3, np.float32)
    for k in range(nb):
        sh = np.zeros(48, np.float32)
        sh[3 * k] = 1.0
        for i, di in enumerate(d):
            dd = np.ascontiguousarray(di, np.float32)
            ob.lib().orc_sh_color(ob.ptr(dd), ob.ptr(sh), order, ob.ptr(col))
            out[i, k] = float(col[0]) - 0.5
    return out


def test_sh_basis_is_orthonormal_by_quadrature():
    # Gauss-Legendre in cos(theta) x uniform phi integrates degree <= 6 products exactly
    nt, nphi = 8, 16
    x, wt = np.polynomial.legendre.leggauss(nt)
    phi = 2 * np.pi * (np.arange(nphi) + 0.5) / nphi
    ct, ph = np.meshgrid(x, phi, indexing="ij")
    st = np.sqrt(1 - ct ** 2)
    d = np.stack([st * np.cos(ph), st * np.sin(ph), ct], -1).reshape(-1, 3)
    w = (wt[:, None] * np.full(nphi, 2 * np.pi / nphi)[None, :]).reshape(-1)
    Y = sh_basis(d, 3)
    G = (Y * w[:, None]).T @ Y
    assert np.allclose(G, np.eye(16), atol=2e-5)


def test_sh_order_zero_special_case():
    sh = np.zeros(48, np.float32)
    sh[:3] = (0.3, -0.7, 1.5)
    d = np.array([0.6, 0.0, 0.8], np.float32)
    col = np.zeros(3, np.float32)
    for order in range(4):
        ob.lib().orc_sh_color(ob.ptr(d), ob.ptr(sh), order, ob.ptr(col))
        want = np.float32(0.5) + np.float32(ob.lib().orc_sh_constant(0)) * sh[:3]
        assert np.array_equal(col, want.astype(np.float32))


def test_exp2_against_libm():
    # the fragment's Gaussian is 2^-(|S tc|^2); exp2 clamps at -125 (far below
    # any alpha of 1/255) and must be exact where 2^y is a power of two
    ys = np.concatenate([np.linspace(-125.0, 0, 250001, dtype=np.float32),
                         np.float32([-0.5, -1e-3, -1e-7, 0.0, -4.0, -0.49999997, -0.50000006])])
    got = np.array([ob.exp2(float(y)) for y in ys[::7]], np.float32)
    want = np.exp2(ys[::7].astype(np.float64)).astype(np.float32)
    ulp = np.abs(got.view(np.int32).astype(np.int64) - want.view(np.int32).astype(np.int64))
    assert ulp.max() <= 4
    for k in range(0, 126):
        assert ob.exp2(-float(k)) == 2.0 ** -k
    assert ob.exp2(-200.0) == ob.exp2(-125.0) == 2.0 ** -125


def test_exp2_gaussian_scale():
    # 2^-((S x)^2) with S = sqrt(log2(e)/2) is e^(-x^2/2)
    S = np.float32(0.84932180)
    assert abs(float(S) ** 2 - math.log2(math.e) / 2) < 1e-7
    for x in (0.0, 0.5, 1.0, 1.7, 2.0, 2.8):
        u = np.float32(S * np.float32(x))
        got = ob.exp2(-float(np.float32(u * u)))
        assert abs(got - math.exp(-x * x / 2)) <= 2e-6 * max(1e-3, math.exp(-x * x / 2)) + 1e-7


def test_serialized_dtype_matches_reference_layout():
    d = SERIALIZED_SPLAT_DTYPE
    assert d.itemsize == 248
    assert [d.fields[k][1] for k in ("center", "n", "color", "alpha", "scale", "rotation")] == \
        [0, 12, 24, 216, 220, 232]
