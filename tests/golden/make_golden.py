"""Writes the golden fixtures of this directory from the oracle.

The reference ships no test vectors, so these pin the oracle (and, through
the GPU tests, the HIP path) against change:

  dev_test.json        scene.rs:157-192 dev-test scene seen from (0, 0, -5):
                       sorted keys and indices, 10-float records
  gauss1k.npz          1000 gaussian splats (seed 1), 320x240, pose
                       sphere_pose(1): keys, indices, records, and the
                       64x64 crop of the unorm8 and f32 images at (128, 88)
  entries_sha256.json  sha256 of the sorted (key, index) entries and the
                       visible count of the 100k / 1280x720 / 60 degree
                       camera-at-distance-5 workload and the 1M / 1920x1080
                       sphere_pose(7) workload

    python tests/golden/make_golden.py      # rewrites the fixtures
"""
from __future__ import annotations

import hashlib
import json
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle_bind as ob  # noqa: E402
from splatter_amd.synth import Workload, look_at, sphere_pose  # noqa: E402

CROP = (128, 88, 64)  # x0, y0, size of the stored image crop


def dev_test():
    s = ob.dev_test_scene()
    cam, view = look_at((0.0, 0.0, -5.0), (0.0, 0.0, 0.0))
    u = ob.uniforms(cam, view, 640, 480)
    keys, idx, rec, _ = ob.render_frame(s, u, sh_order=0, covariance=False, with_image=False)
    return {"camera": [0.0, 0.0, -5.0], "size": [640, 480], "sh_order": 0, "covariance": False,
            "keys": [int(k) for k in keys], "indices": [int(i) for i in idx],
            "records": [[float(x) for x in r] for r in rec]}


def gauss1k():
    splats = ob.decode(Workload("gaussian", 1000, seed=1).all())
    cam, view = sphere_pose(1)
    u = ob.uniforms(cam, view, 320, 240)
    keys, idx, rec, img = ob.render_frame(splats, u, sh_order=3, covariance=True)
    _, _, _, img32 = ob.render_frame(splats, u, sh_order=3, covariance=True, blend_f32=True)
    x0, y0, n = CROP
    return {"keys": keys, "indices": idx, "records": rec,
            "image_unorm8": img[y0:y0 + n, x0:x0 + n].copy(), "image_f32": img32[y0:y0 + n, x0:x0 + n].copy()}


def entries_hash(workload, size, pose, fov):
    splats = workload.all()
    cam, view = pose
    u = ob.uniforms(cam, view, size[0], size[1], fov=fov)
    keys, idx = ob.cpu_depth_sort(np.ascontiguousarray(splats["center"]), ob.vp_of(u), 1.1)
    h = hashlib.sha256(keys.astype("<u4").tobytes() + idx.astype("<u4").tobytes()).hexdigest()
    return {"visible": int(len(idx)), "sha256": h}


def entries():
    return {
        "gaussian100k_1280x720_fov60_distance5": entries_hash(
            Workload("gaussian", 100_000, seed=0), (1280, 720), look_at((0.0, 0.0, -5.0), (0.0, 0.0, 0.0)),
            math.radians(60)),
        "gaussian1m_1920x1080_fov90_sphere_pose7": entries_hash(
            Workload("gaussian", 1_000_000, seed=1), (1920, 1080), sphere_pose(7), math.pi / 2),
    }


def main():
    with open(os.path.join(HERE, "dev_test.json"), "w") as f:
        json.dump(dev_test(), f, indent=1)
    np.savez_compressed(os.path.join(HERE, "gauss1k.npz"), **gauss1k())
    with open(os.path.join(HERE, "entries_sha256.json"), "w") as f:
        json.dump(entries(), f, indent=1)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
