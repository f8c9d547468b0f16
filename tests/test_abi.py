"""The C-ABI library loads and exports every entry point include/splatter.h
declares (no device calls: runs without a GPU)."""
import ctypes
import os
import re

import numpy as np

import splatter_amd
from splatter_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "splatter.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(spl_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.lib()
    names = declared_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
    bound = {s[0] for s in _lib.SIGNATURES}
    assert set(names) == bound


def test_configuration_layout_and_defaults():
    c = _lib.SplConfiguration()
    _lib.lib().spl_configuration_default(ctypes.byref(c))
    py = splatter_amd.Configuration().to_c()
    for name, _ in _lib.SplConfiguration._fields_:
        assert getattr(c, name) == getattr(py, name), name
    assert c.max_splat_count == 4 * 1024 * 1024 and c.radix_bits_per_digit == 8
    assert c.spherical_harmonics_order == 2 and abs(c.field_of_view_y - np.pi / 2) < 1e-7


def test_error_path_without_device():
    """spl_parse_file_header needs no device: a missing file is an IO error."""
    with np.testing.assert_raises(splatter_amd.SplatterError):
        splatter_amd.Scene.parse_file_header("/nonexistent/file.ply")


def test_parse_file_header(tmp_path):
    header = b"ply\nformat binary_little_endian 1.0\nelement vertex 3\nproperty float x\nend_header\n"
    p = tmp_path / "s.ply"
    p.write_bytes(header + b"\0" * 3 * 248)
    hs, n = splatter_amd.Scene.parse_file_header(str(p))
    assert (hs, n) == (len(header), 3)
