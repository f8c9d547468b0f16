"""The staged autotuner's selection rule (host logic, no device): each stage
tries its candidates in order of preference and a later candidate replaces
the pick only when it is more than the tie margin faster."""
from types import SimpleNamespace

from splatter_amd import autotune as at


def _run(monkeypatch, fps_of):
    order = []

    def fake_measure(sa, device, base, scene_loader, poses, width, height, cand, steps, in_flight):
        order.append(cand.key())
        return fps_of(cand)

    monkeypatch.setattr(at, "measure", fake_measure)
    base = SimpleNamespace(raster_batch=128)
    res = at.autotune_staged(None, None, base, None, None, 64, 48, 16, steps=1, in_flight=2)
    return res, order


def test_preference_order_and_tie(monkeypatch):
    # 8192-key tiles 2% ahead of 16384: inside the 3% margin, so 16384 stays
    def fps_of(c):
        f = {16384: 1500.0, 8192: 1530.0, 4096: 1400.0}[c.sort_tile_keys]
        f += 1.0 if c.radix_bits_per_digit == 10 else 0.0
        f -= 60.0 * c.raster_longest_first + 100.0 * (c.raster_batch == 256)
        return f
    res, order = _run(monkeypatch, fps_of)
    assert order[0] == "r10-t16384-b128-m3-o0-v16" and order[1] == "r8-t16384-b128-m3-o0-v16"
    assert res.best.sort_tile_keys == 16384 and res.best.radix_bits_per_digit == 10
    assert res.best.raster_batch == 128 and res.best.raster_longest_first == 0
    assert res.best.raster_mode == 3  # the preferred walk, no slower than entry-major here
    assert len(res.table) == 6 + 8


def test_clear_win_replaces_the_pick(monkeypatch):
    # 4096-key tiles 10% ahead: a clear win; longest-first 5% ahead: taken too
    def fps_of(c):
        f = {16384: 1000.0, 8192: 1010.0, 4096: 1100.0}[c.sort_tile_keys]
        return f * (1.05 if c.raster_longest_first else 1.0)
    res, _ = _run(monkeypatch, fps_of)
    assert res.best.sort_tile_keys == 4096 and res.best.raster_longest_first == 1
