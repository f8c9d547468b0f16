"""Synthetic scenes and camera poses for the measured workloads.

Splats are generated as 248-byte SerializedSplat records (scene.rs:17-24) and
go through the same decode as a PLY file (scene.rs:225-249, done by
Scene.load_records).  Generation is chunked with per-chunk seeds so that a
32M-splat scene streams through host memory.

Workloads (sizes as the names say):
  gaussian      positions N(0, 2^2), log-scale N(-4, 0.5), rotation N(0,1)^4,
                opacity logit N(0, 1.5), SH DC N(0, 0.5), higher bands N(0, 0.05)
  clustered     as gaussian, but 80% of positions from 64 anisotropic Gaussian
                clusters (centres U[-10,10]^3, axis sigmas log-uniform in
                [0.1, 1.0], random rotations) and 20% on a ground slab
                x,y in [-30,30], z in [-0.2,0.2]
  tight         positions from 16 isotropic clusters, sigma 0.05, centres
                U[-5,5]^3: dense, tie-heavy depths
"""
from __future__ import annotations

import math

import numpy as np

from . import SERIALIZED_SPLAT_DTYPE

CHUNK = 1 << 20


def _attributes(rng: np.random.Generator, n: int) -> np.ndarray:
    rec = np.zeros(n, dtype=SERIALIZED_SPLAT_DTYPE)
    rec["scale"] = rng.normal(-4.0, 0.5, size=(n, 3)).astype(np.float32)
    rec["rotation"] = rng.normal(0.0, 1.0, size=(n, 4)).astype(np.float32)
    rec["alpha"] = rng.normal(0.0, 1.5, size=n).astype(np.float32)
    color = rng.normal(0.0, 0.05, size=(n, 48)).astype(np.float32)
    color[:, :3] = rng.normal(0.0, 0.5, size=(n, 3)).astype(np.float32)
    rec["color"] = color
    return rec


def _random_rotations(rng: np.random.Generator, k: int) -> np.ndarray:
    q = rng.normal(size=(k, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
        np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], -1),
        np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1),
    ], 1)


class Workload:
    """Deterministic splat generator: chunk(i) is reproducible on its own."""

    def __init__(self, kind: str, count: int, seed: int = 0):
        if kind not in ("gaussian", "clustered", "tight"):
            raise ValueError(kind)
        self.kind, self.count, self.seed = kind, int(count), int(seed)
        rng = np.random.default_rng([self.seed, 0x5EED])
        if kind == "clustered":
            self.k = 64
            self.centres = rng.uniform(-10.0, 10.0, size=(self.k, 3))
            self.sigmas = np.exp(rng.uniform(math.log(0.1), math.log(1.0), size=(self.k, 3)))
            self.rots = _random_rotations(rng, self.k)
        elif kind == "tight":
            self.k = 16
            self.centres = rng.uniform(-5.0, 5.0, size=(self.k, 3))

    def chunk(self, i: int) -> np.ndarray:
        lo = i * CHUNK
        n = min(CHUNK, self.count - lo)
        rng = np.random.default_rng([self.seed, i])
        rec = _attributes(rng, n)
        if self.kind == "gaussian":
            pos = rng.normal(0.0, 2.0, size=(n, 3))
        elif self.kind == "clustered":
            pos = np.empty((n, 3))
            in_cluster = rng.uniform(size=n) < 0.8
            m = int(in_cluster.sum())
            cid = rng.integers(0, self.k, size=m)
            local = rng.normal(size=(m, 3)) * self.sigmas[cid]
            pos[in_cluster] = self.centres[cid] + np.einsum("nij,nj->ni", self.rots[cid], local)
            s = n - m
            pos[~in_cluster] = np.stack([rng.uniform(-30, 30, s), rng.uniform(-30, 30, s),
                                         rng.uniform(-0.2, 0.2, s)], -1)
        else:
            cid = rng.integers(0, self.k, size=n)
            pos = self.centres[cid] + rng.normal(size=(n, 3)) * 0.05
        rec["center"] = pos.astype(np.float32)
        return rec

    def chunks(self):
        for i in range((self.count + CHUNK - 1) // CHUNK):
            yield i * CHUNK, self.chunk(i)

    def all(self) -> np.ndarray:
        return np.concatenate([c for _, c in self.chunks()]) if self.count else \
            np.zeros(0, SERIALIZED_SPLAT_DTYPE)


def load_scene(scene, workload: Workload) -> None:
    """Streams the workload into a Scene through Scene.load_records."""
    for first, rec in workload.chunks():
        scene.load_records(first, rec)


def look_at(eye, target, up=(0.0, 1.0, 0.0)):
    """Rigid camera at `eye` looking at `target`: view space +x right, +y up,
    +z forward (the reference's projection puts depth on +z, utils.rs:45-55).
    Returns (camera_matrix, view_matrix) as float32 (4, 4) column-major
    arrays, i.e. m[c] is column c, matching motor3d_to_mat4 (utils.rs:31-42)."""
    e = np.asarray(eye, np.float64)
    f = np.asarray(target, np.float64) - e
    f /= np.linalg.norm(f)
    u = np.asarray(up, np.float64)
    if abs(np.dot(u, f)) > 0.999:
        u = np.array([0.0, 0.0, 1.0]) if abs(f[2]) < 0.9 else np.array([1.0, 0.0, 0.0])
    r = np.cross(u, f)
    r /= np.linalg.norm(r)
    u = np.cross(f, r)
    cam = np.zeros((4, 4))
    cam[0, :3], cam[1, :3], cam[2, :3], cam[3, :3] = r, u, f, e
    cam[3, 3] = 1.0
    R = np.stack([r, u, f])           # rows: world -> view rotation
    view = np.zeros((4, 4))
    for c in range(3):
        view[c, :3] = R[:, c]
    view[3, :3] = -R @ e
    view[3, 3] = 1.0
    return cam.astype(np.float32), view.astype(np.float32)


def sphere_pose(seed: int, radius: float = 6.0):
    rng = np.random.default_rng([seed, 0xCA3])
    v = rng.normal(size=3)
    v /= np.linalg.norm(v)
    return look_at(radius * v, (0.0, 0.0, 0.0), (0.0, 1.0, 0.0))


def orbit_poses(count: int, radius: float = 6.0, elevation_deg: float = 20.0):
    """`count` poses at equal azimuth steps around the z axis (z up)."""
    el = math.radians(elevation_deg)
    out = []
    for k in range(count):
        az = 2.0 * math.pi * k / count
        eye = (radius * math.cos(el) * math.cos(az), radius * math.cos(el) * math.sin(az),
               radius * math.sin(el))
        out.append(look_at(eye, (0.0, 0.0, 0.0), (0.0, 0.0, 1.0)))
    return out
