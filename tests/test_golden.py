"""The oracle reproduces the committed golden fixtures (tests/golden/):
dev-test scene entries and records, a 1000-splat frame's entries, records
and image crops, and the sha256 of two full-size sorted entry lists."""
import json
import os

import numpy as np

import golden.make_golden as mg

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_dev_test_fixture():
    want = json.load(open(os.path.join(HERE, "dev_test.json")))
    got = mg.dev_test()
    assert got["keys"] == want["keys"] and got["indices"] == want["indices"] == [0, 1, 2]
    # axis-aligned splats hit the contour path's 0/0 (shaders.wgsl:202): NaN rectangle
    assert np.array_equal(np.float32(got["records"]), np.float32(want["records"]), equal_nan=True)


def test_gauss1k_fixture():
    want = np.load(os.path.join(HERE, "gauss1k.npz"))
    got = mg.gauss1k()
    for k in ("keys", "indices", "records", "image_unorm8", "image_f32"):
        assert np.array_equal(got[k], want[k]), k
    assert want["image_unorm8"].any()


def test_entry_hashes_fixture():
    want = json.load(open(os.path.join(HERE, "entries_sha256.json")))
    assert mg.entries() == want
