"""The HIP path against the committed golden fixtures (tests/golden/): the
same workloads rendered through the C ABI give the fixtures' sorted entries
bit for bit, their records within 1e-5 and their image crops within 2/255."""
import hashlib
import json
import math
import os

import numpy as np
import pytest

import splatter_amd as sa
from gpu_common import config, device, image_diff, normwise_rel, say
from splatter_amd.synth import Workload, look_at, sphere_pose

pytestmark = pytest.mark.gpu
HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CROP = (128, 88, 64)


def render(records_or_none, n, w, h, pose, cfg, dev_test=False):
    r = sa.Renderer(device(), cfg)
    scene = sa.Scene.new_dev_test(r) if dev_test else sa.Scene(r, n)
    if not dev_test:
        scene.load_records(0, records_or_none)
    f = sa.Frame(device(), w, h)
    r.render_frame(f, pose[0], pose[1], scene)
    img = f.read()[..., [2, 1, 0, 3]]
    keys, idx = r.read_entries()
    return r, keys, idx, img


def test_dev_test_fixture():
    want = json.load(open(os.path.join(HERE, "dev_test.json")))
    cfg = config(spherical_harmonics_order=0, use_covariance_for_scale=False, max_splat_count=16)
    r, keys, idx, _ = render(None, 3, 640, 480, look_at((0.0, 0.0, -5.0), (0.0, 0.0, 0.0)), cfg, dev_test=True)
    assert list(keys) == want["keys"] and list(idx) == want["indices"]
    rec = r.read_records()
    assert np.array_equal(rec, np.float32(want["records"]), equal_nan=True)


@pytest.mark.parametrize("blend", ["unorm8", "f32"])
def test_gauss1k_fixture(blend):
    want = np.load(os.path.join(HERE, "gauss1k.npz"))
    cfg = config(max_splat_count=4096, blend_mode=sa.BlendMode.F32 if blend == "f32" else sa.BlendMode.UNORM8)
    r, keys, idx, img = render(Workload("gaussian", 1000, seed=1).all(), 1000, 320, 240, sphere_pose(1), cfg)
    assert np.array_equal(keys, want["keys"]) and np.array_equal(idx, want["indices"])
    rec = r.read_records()
    assert np.array_equal(rec[:, :4], want["records"][:, :4])
    assert normwise_rel(rec[:, 4:], want["records"][:, 4:]).max() <= 1e-5
    x0, y0, n = CROP
    d = image_diff(img[y0:y0 + n, x0:x0 + n], want[f"image_{blend}"])
    say(f"gauss1k {blend}: crop max diff {d}/255")
    assert d <= 2


@pytest.mark.parametrize("name", ["gaussian100k_1280x720_fov60_distance5", "gaussian1m_1920x1080_fov90_sphere_pose7"])
def test_entry_hash_fixture(name):
    want = json.load(open(os.path.join(HERE, "entries_sha256.json")))[name]
    if name.startswith("gaussian100k"):
        wl, w, h, pose, fov = Workload("gaussian", 100_000, seed=0), 1280, 720, \
            look_at((0.0, 0.0, -5.0), (0.0, 0.0, 0.0)), math.radians(60)
    else:
        wl, w, h, pose, fov = Workload("gaussian", 1_000_000, seed=1), 1920, 1080, sphere_pose(7), math.pi / 2
    cfg = config(max_splat_count=wl.count, field_of_view_y=fov)
    _, keys, idx, _ = render(wl.all(), wl.count, w, h, pose, cfg)
    got = hashlib.sha256(keys.astype("<u4").tobytes() + idx.astype("<u4").tobytes()).hexdigest()
    say(f"{name}: visible {len(idx)} (fixture {want['visible']})")
    assert len(idx) == want["visible"] and got == want["sha256"]
