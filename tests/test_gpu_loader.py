"""Scene upload paths through the C ABI.  Scene::load_chunk (scene.rs:215-249)
reads the records of a binary PLY file after its header; Scene.load_records
takes the same 248-byte records from memory.  Both decode on host threads into
pinned staging and upload 2^18 records at a time, two chunks in flight, so the
ranges below cross chunk and thread boundaries.  Every path must give the scene
that the oracle decodes from the same records."""
import numpy as np
import pytest

import splatter_amd as sa
from gpu_common import Case, config, image_diff, say
from splatter_amd.synth import Workload, sphere_pose

pytestmark = pytest.mark.gpu


def _write_ply(path, records):
    header = (b"ply\nformat binary_little_endian 1.0\ncomment written by the loader test\n"
              + f"element vertex {len(records)}\n".encode()
              + b"".join(f"property float f{i}\n".encode() for i in range(62)) + b"end_header\n")
    with open(path, "wb") as f:
        f.write(header)
        f.write(records.tobytes())
    return len(header)


def test_load_chunk_and_ragged_load_records(tmp_path):
    w = Workload("gaussian", 600_000, seed=13)
    case = Case(w, config(), 640, 480, sphere_pose(13, radius=4))
    say("oracle 600k")
    _, _, _, want = case.oracle()
    ref = case.render()
    assert image_diff(ref, want) <= 2

    path = tmp_path / "scene.ply"
    written = _write_ply(path, case.records)
    hs, n = sa.Scene.parse_file_header(str(path))
    assert (hs, n) == (written, w.count)

    say("load_chunk in three ranges, the last past the end")
    from_file = sa.Scene(case.renderer, n)
    for b, e in ((0, 1), (1, 270_001), (270_001, n + 5000)):
        from_file.load_chunk(str(path), hs, b, e)
    say("load_records in ragged pieces")
    pieces = sa.Scene(case.renderer, n)
    for b, e in ((0, 4097), (4097, 4098), (4098, 530_000), (530_000, n)):
        pieces.load_records(b, case.records[b:e])

    loaded = case.scene
    try:
        for scene in (from_file, pieces):
            case.scene = scene
            img = case.render()
            assert np.array_equal(img, ref)
    finally:
        case.scene = loaded
        from_file.close()
        pieces.close()


def test_load_chunk_errors(tmp_path):
    w = Workload("gaussian", 1000, seed=14)
    r = sa.Renderer(sa.Device(0), config(max_splat_count=1000))
    s = sa.Scene(r, 1000)
    path = tmp_path / "short.ply"
    hs = _write_ply(path, w.all()[:600])
    with pytest.raises(sa.SplatterError):
        s.load_chunk(str(path), hs, 0, 1000)  # the file holds 600 records
    with pytest.raises(sa.SplatterError):
        s.load_chunk(str(tmp_path / "missing.ply"), hs, 0, 10)
    s.load_chunk(str(path), hs, 500, 500)  # empty range: nothing read
    with pytest.raises(sa.SplatterError):
        s.load_records(900, w.all()[:200])  # past splat_count
    s.close()
    r.close()
