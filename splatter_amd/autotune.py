"""Autotuner for the batched form: picks the radix digit width, the radix-sort
tile (keys per tile, i.e. the sort block's lane count), the rasteriser's
batch (entries a tile block stages per round), walk (entry-major, per-pixel,
per 4x4 sub-block or per-lane over coverage masks) and tile order, and the batch width (views per launch sequence)
for one scene, resolution and pose set by timing the candidates on the
device, and can keep the answer in a JSON cache.

The candidates all render identical images and sort orders (the tests check
each option against the oracle), so the tuner is free to pick on speed alone.
"""
from __future__ import annotations

import itertools
import json
import os
import time
from dataclasses import dataclass, field

import numpy as np

RADIX_BITS = (8, 10)
SORT_TILE_KEYS = (4096, 8192, 16384)
RASTER_MODE = (0, 1, 2, 3)
RASTER_BATCH = (128, 256)
RASTER_LONGEST_FIRST = (0, 1)
BATCH_VIEWS = (8, 16, 32)


@dataclass
class Candidate:
    radix_bits_per_digit: int
    sort_tile_keys: int
    max_views: int
    raster_mode: int = 0
    raster_batch: int = 128
    raster_longest_first: int = 0

    def key(self) -> str:
        return (f"r{self.radix_bits_per_digit}-t{self.sort_tile_keys}-b{self.raster_batch}-m{self.raster_mode}"
                f"-o{self.raster_longest_first}-v{self.max_views}")

    def replace(self, **kw) -> "Candidate":
        d = dict(self.__dict__)
        d.update(kw)
        return Candidate(**d)


@dataclass
class TuneResult:
    best: Candidate
    frames_per_s: float
    table: list = field(default_factory=list)  # (candidate key, frames/s)

    def apply(self, config) -> None:
        config.radix_bits_per_digit = self.best.radix_bits_per_digit
        config.sort_tile_keys = self.best.sort_tile_keys
        config.max_views = self.best.max_views
        config.raster_mode = self.best.raster_mode
        config.raster_batch = self.best.raster_batch
        config.raster_longest_first = self.best.raster_longest_first

    def to_json(self) -> dict:
        return {"best": self.best.__dict__, "frames_per_s": self.frames_per_s, "table": self.table}

    @classmethod
    def from_json(cls, d: dict) -> "TuneResult":
        return cls(Candidate(**d["best"]), float(d["frames_per_s"]), [tuple(x) for x in d.get("table", [])])


def candidates(radix_bits=RADIX_BITS, sort_tile_keys=SORT_TILE_KEYS, batch_views=BATCH_VIEWS,
               raster_mode=RASTER_MODE):
    return [Candidate(r, t, v, w) for r, t, v, w in
            itertools.product(radix_bits, sort_tile_keys, batch_views, raster_mode)]


def measure(sa, device, base, scene_loader, poses, width: int, height: int, cand: Candidate,
            steps: int = 3, in_flight: int = 1, reps: int = 2) -> float:
    """frames/s of `steps` graph-replayed batches of cand.max_views views
    (the best of `reps` timed runs, which damps the run-to-run noise that
    otherwise decides between candidates a few percent apart),
    with `in_flight` renderers taking alternate batches, each with its own
    frames.  One renderer runs on the device stream.  With more than one,
    each runs on a stream of its own: batches on the device stream did not
    overlap with those of an own-stream renderer, measured on MI355X."""
    cfg = type(base)(**{k: getattr(base, k) for k in base.__dataclass_fields__})
    cfg.radix_bits_per_digit = cand.radix_bits_per_digit
    cfg.sort_tile_keys = cand.sort_tile_keys
    cfg.max_views = cand.max_views
    cfg.raster_mode = cand.raster_mode
    cfg.raster_batch = cand.raster_batch
    cfg.raster_longest_first = cand.raster_longest_first
    rs = ([sa.Renderer(device, cfg)] if in_flight <= 1 else
          [sa.Renderer(device, cfg, own_stream=True) for _ in range(in_flight)])
    try:
        scene = scene_loader(rs[0])
        nv = cand.max_views
        fsets = [[sa.Frame(device, width, height) for _ in range(nv)] for _ in rs]

        def batch(k):
            j = k % len(rs)
            sel = [poses[(k * nv + i) % len(poses)] for i in range(nv)]
            rs[j].render_views(fsets[j], [p[0] for p in sel], [p[1] for p in sel], scene)

        def sync():
            device.synchronize()
            for r in rs:
                if r.own_stream:
                    r.synchronize()

        for k in range(len(rs)):
            batch(k)
        sync()
        best = 0.0
        for rep in range(max(1, reps)):
            t0 = time.perf_counter()
            for k in range(steps):
                batch(k + len(rs) + rep * steps)
            sync()
            best = max(best, steps * nv / (time.perf_counter() - t0))
        return best
    finally:
        for r in rs:
            r.close()


def autotune(sa, device, base, scene_loader, poses, width: int, height: int, cands=None, steps: int = 3,
             log=None) -> TuneResult:
    """Times every candidate; scene_loader(renderer) returns the scene to
    render (it may return the same resident scene every time)."""
    cands = cands or candidates()
    table = []
    for c in cands:
        fps = measure(sa, device, base, scene_loader, poses, width, height, c, steps)
        table.append((c.key(), round(fps, 2)))
        if log:
            log(f"autotune {c.key()}: {fps:.1f} frames/s")
    best_i = int(np.argmax([t[1] for t in table]))
    return TuneResult(cands[best_i], table[best_i][1], table)


def cache_key(device_name: str, splats: int, width: int, height: int, sh_order: int) -> str:
    bucket = 1 << max(0, int(splats - 1).bit_length())
    return f"{device_name}|n<={bucket}|{width}x{height}|sh{sh_order}"


def load_cached(path: str, key: str):
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return TuneResult.from_json(d[key]) if key in d else None


def store_cached(path: str, key: str, result: TuneResult) -> None:
    d = {}
    if os.path.exists(path):
        with open(path) as f:
            d = json.load(f)
    d[key] = result.to_json()
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as f:
        json.dump(d, f, indent=1, sort_keys=True)


# A candidate replaces the current pick only if it is this much faster.  The
# in-process timings of one candidate scatter by 2-3% on MI355X, and they
# once put 8192-key sort tiles 2.2% ahead of 16384-key ones that fresh
# processes then measured 3-5% faster.
TIE = 0.03


def autotune_staged(sa, device, base, scene_loader, poses, width: int, height: int, views: int, steps: int = 3,
                    log=None, walk=(3, 0), in_flight: int = 1, tie: float = TIE) -> TuneResult:
    """Coordinate search: the sort's digit width x tile with the raster at
    its defaults, then the raster's batch x tile order (x walk) with the best
    sort, at a fixed batch width and number of batches in flight.  Each stage
    tries its candidates in order of preference (larger sort tiles, which
    pay their look-back over more keys, and the raster's default form
    first); a later one is picked only if it beats the pick by more than
    `tie`."""
    table = []

    def run(cands):
        pick = None
        for c in cands:
            fps = measure(sa, device, base, scene_loader, poses, width, height, c, steps, in_flight)
            table.append((c.key(), round(fps, 2)))
            if log:
                log(f"autotune {c.key()}: {fps:.1f} frames/s")
            if pick is None or fps > pick[1] * (1.0 + tie):
                pick = (c, fps)
        return pick

    start = Candidate(8, 4096, views, walk[0], base.raster_batch, 0)
    c, _ = run([start.replace(radix_bits_per_digit=r, sort_tile_keys=t)
                for t in sorted(SORT_TILE_KEYS, reverse=True) for r in sorted(RADIX_BITS, reverse=True)])
    c, fps = run([c.replace(raster_batch=b, raster_longest_first=o, raster_mode=w)
                  for b in RASTER_BATCH for o in RASTER_LONGEST_FIRST for w in walk])
    return TuneResult(c, round(fps, 2), table)
