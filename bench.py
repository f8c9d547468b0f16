"""Benchmark of the splat pipeline on one MI355X.

Workload: 6M clustered Gaussians (64 anisotropic clusters with 80% of the
splats + a 20% ground slab), SH degree 3, 1920x1080, a 256-pose orbit
(radius 6, elevation 20 deg), 16 views per batched launch.  One step = one
batch of 16 views.  The batched loop keeps two batches in flight (two
renderers with their own frames, alternating, as the serving form's slots
do), so one batch's raster overlaps the next one's sort and projection; the
plain path renders one view at a time on one renderer.  Output: one JSON
line (see the keys below).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--in-flight 2]

Per-stage device times come from HIP events that the library records on its
own stream around every stage of every frame of a second pass over the same
batches on one renderer (one batch at a time, so each stage's events
bracket it alone).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "6M-splat 1080p frames/s batched vs plain; sort Gkeys/s; p50/p99 render latency"
HBM_PEAK_GBS = 8000.0
# VALU issue peak: 256 CUs x 4 SIMDs x 2.4 GHz, one wave64 instruction per SIMD every 4 cycles
VALU_PEAK_PER_S = 256 * 4 * 2.4e9 / 4
STAGES = ["cull_key_compact", "digit_histograms", "depth_sort", "project", "bin", "pair_sort",
          "tile_ranges", "raster"]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def stage_bytes(V: float, K: float, N: int, W: int, H: int, L: int, P: int, Q: int, nv: int) -> dict:
    """Algorithmic HBM bytes per frame of each stage (what each kernel must
    move at minimum for its data layout).  Reads shared by the nv views of a
    batch (centres, visibility masks, splat attributes) are charged 1/nv per
    frame; splat attributes are charged for V splats per view (a lower bound
    on the splats some view of the batch sees)."""
    sh = 12 * (L + 1) ** 2
    return {
        "cull_key_compact": (12 * N + 4 * N) / nv + 8 * V,   # centres + visibility out; key + index out
        "digit_histograms": 4 * V,                          # keys in
        "depth_sort": 16 * P * V,                           # key + value in and out per pass
        "project": 4 * N / nv + (12 + 32 + sh) * V / nv + 48 * V,  # masks, attributes in; record out
        "bin": (4 + 16) * V + 8 * K,                        # order + record box in; (tile, record) pairs out
        "pair_sort": 16 * Q * K,
        "tile_ranges": 4 * K,
        "raster": (4 + 48) * K + 4 * W * H,                 # pair value + record gather; pixels out
    }


STAGE_KERNELS = {"cull_key_compact": "k_cull_views", "digit_histograms": "k_digit_histograms",
                 "depth_sort": "k_onesweep_pass", "project": "k_project", "bin": "k_bin",
                 "pair_sort": "k_onesweep_pass", "tile_ranges": "k_tile_ranges", "raster": "k_raster"}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_6M_clustered_1080p_16views.txt")
KERNELS_FILE = os.path.join(ROOT, "profiles", "kernels_6M_clustered_1080p_16views.md")


def measured_valu(stage: str):
    """(kernel, SQ_INSTS_VALU per launch) of the stage's kernel from the
    committed rocprofv3 counter summary of this workload, or None."""
    if not os.path.exists(KERNELS_FILE):
        return None
    want, cols, in_sq = STAGE_KERNELS.get(stage), None, False
    for line in open(KERNELS_FILE):
        if line.startswith("## "):
            in_sq = line.startswith("## SQ counters")
            continue
        f = line.split()
        if not in_sq or not f:
            continue
        if f[0] == "kernel":
            cols = f
        elif cols and "INSTS_VALU" in cols and f[0].startswith(want):
            return f[0], float(f[cols.index("INSTS_VALU")])
    return None


def measured_traffic(stage: str):
    """HBM bytes per launch of the stage's kernel (FETCH_SIZE + WRITE_SIZE,
    scaled against a device copy of known size) from the committed rocprofv3
    summary of this workload, or None when there is none."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    want = STAGE_KERNELS.get(stage)
    for line in open(TRAFFIC_FILE):
        f = line.split()
        if len(f) >= 4 and not line.startswith("#") and f[0].startswith(want) and f[0] != "kernel":
            return int(float(f[2]) + float(f[3]))
    return None


def host_cpu() -> dict:
    """Host cores visible to this process (nproc) and the CPU model."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": len(os.sched_getaffinity(0)), "cpu_model": model}


def cpu_baseline(positions: np.ndarray, poses, W: int, H: int, budget_s: float = 12.0) -> dict:
    """The reference's CPU depth-sort fallback (renderer.rs:372-396) restated
    in C, on the same scene and orbit views: median seconds per view on one
    core (the reference's single-threaded iterator and sort_by), then on all
    host cores this process may use, at most 16 (parallel transform and
    parallel stable merge sort, checked to give the one-core entries); and
    the oracle's full CPU frame (projection, rectangles, raster, unorm8
    blend) of the 100k-Gaussian 1280x720 60-degree view, median of 10 runs
    after 2 warm-ups."""
    import ctypes
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind as ob
    from splatter_amd.synth import Workload, look_at
    host = host_cpu()
    threads = max(1, min(16, host["nproc"]))
    pos = np.ascontiguousarray(positions, np.float32)
    t1, tn, vis = [], [], []
    start = time.perf_counter()
    for cam, view in poses:
        vp = np.ascontiguousarray(ob.vp_of(ob.uniforms(cam, view, W, H)), np.float32)
        vc = ctypes.c_size_t()
        t1.append(ob.lib().orc_time_cpu_depth_sort(ob.ptr(pos), len(pos), ob.ptr(vp), 1.1, 1, ctypes.byref(vc)))
        vis.append(vc.value)
        tn.append(ob.lib().orc_time_cpu_depth_sort_threads(ob.ptr(pos), len(pos), ob.ptr(vp), 1.1, threads, 3,
                                                           ctypes.byref(vc)))
        if time.perf_counter() - start > budget_s:
            break
    # the all-cores leg gives the one-core leg's entries, bit for bit
    cam, view = poses[0]
    vp = ob.vp_of(ob.uniforms(cam, view, W, H))
    k1, i1 = ob.cpu_depth_sort(pos, vp, 1.1)
    kn, i_n = ob.cpu_depth_sort_threads(pos, vp, 1.1, threads)
    identical = bool(np.array_equal(k1, kn) and np.array_equal(i1, i_n))
    # full CPU frame of the small workload
    sp = ob.decode(Workload("gaussian", 100_000, seed=0).all())
    cam, view = look_at((0, 0, -5), (0, 0, 0))
    u = ob.uniforms(cam, view, 1280, 720, fov=math.radians(60))
    frame_t = []
    for i in range(12):
        t0 = time.perf_counter()
        ob.render_frame(sp, u, 3, True)
        if i >= 2:
            frame_t.append(time.perf_counter() - t0)
    m1, mn = float(np.median(t1)), float(np.median(tn))
    return {"value": round(1.0 / m1, 4), "unit": "frames/s", "cores": 1, "kind": "port",
            "sample": f"cull+key+stable sort (renderer.rs:372-396, C restatement) of the "
                      f"{len(positions)}-splat scene for {len(t1)} orbit views, median {m1 * 1e3:.1f} ms/view "
                      f"on 1 thread, {int(np.mean(vis))} visible",
            "cores_all": threads, "value_all_cores": round(1.0 / mn, 4), "ms_per_view_all_cores": round(mn * 1e3, 2),
            "all_cores_identical": identical, **host,
            "config0_frame_ms": round(float(np.median(frame_t)) * 1e3, 1),
            "config0_frame": "oracle full frame (FP64 projection, SH3, rectangles, raster, unorm8 blend), "
                             "100k Gaussians, 1280x720, 60 deg, camera at distance 5, 1 thread, median of 10"}


def sort_stress(sa, dev, log, n: int = 32_000_000, width: int = 3840, height: int = 2160, views: int = 8) -> dict:
    """The depth sort on the stress workload: 32M splats in 16 tight
    clusters (sigma 0.05, dense and tie-heavy f32 depths), 8 orbit views at
    3840x2160 in one batch.  Per-frame stage times from HIP events around
    every stage of 4 batches, for the two sort shapes the tuner weighs
    (10-bit digits in 16384-key tiles, 8-bit digits in 4096-key tiles);
    Gkeys/s = visible entries / depth-sort time, GB/s = 16 B x passes x
    visible / depth-sort time (key and value read and written per pass)."""
    from splatter_amd.synth import Workload, orbit_poses
    t0 = time.perf_counter()
    shapes = [(10, 16384), (8, 4096)]
    rs = [sa.Renderer(dev, sa.Configuration(spherical_harmonics_order=3, max_splat_count=n, max_views=views,
                                            tile_pairs_per_splat=8, radix_bits_per_digit=b, sort_tile_keys=t))
          for b, t in shapes]
    scene = sa.Scene(rs[0], n)
    for first, rec in Workload("tight", n, seed=4).chunks():
        scene.load_records(first, rec)
    dev.synchronize()
    log(f"sort stress: {n} splats loaded in {time.perf_counter() - t0:.1f}s")
    poses = orbit_poses(views, radius=9.0)
    frames = [sa.Frame(dev, width, height) for _ in range(views)]
    out = {"workload": f"{n // 1_000_000}M splats in 16 tight clusters (sigma 0.05), {width}x{height}, "
                       f"{views} views per batch"}
    for (bits, tile), r in zip(shapes, rs):
        r.render_views(frames, [p[0] for p in poses], [p[1] for p in poses], scene)
        r.enable_timing(True)
        for _ in range(4):
            r.render_views(frames, [p[0] for p in poses], [p[1] for p in poses], scene)
        st = r.stage_times()
        r.enable_timing(False)
        vis = []
        for v in range(views):
            r.select_view(v)
            vis.append(r.visible_count())
        r.select_view(0)
        V = float(np.mean(vis))
        P = math.ceil(30 / bits)
        sort_ms = float(st[STAGES.index("depth_sort")])
        pair_ms = float(st[STAGES.index("pair_sort")])
        gbs = 16.0 * P * V / (sort_ms * 1e-3) / 1e9
        out[f"r{bits}_t{tile}"] = {"visible_per_view": int(V), "depth_sort_ms": round(sort_ms, 4),
                                   "gkeys_per_s": round(V / (sort_ms * 1e-3) / 1e9, 2), "passes": P,
                                   "GB/s": round(gbs, 1), "frac_of_8TBs": round(gbs / HBM_PEAK_GBS, 4),
                                   "pair_sort_ms": round(pair_ms, 4)}
        log(f"sort stress r{bits} t{tile}: {out[f'r{bits}_t{tile}']}")
    for o in frames + [scene] + rs:
        o.close()
    return out


def config0_gpu_ms(sa, dev, repeats: int = 50) -> float:
    """The plain GPU path on the small workload (100k Gaussians, 1280x720,
    60 degrees, camera at distance 5): median submit-to-ready ms per frame."""
    from splatter_amd.synth import Workload, look_at
    w = Workload("gaussian", 100_000, seed=0)
    cfg = sa.Configuration(spherical_harmonics_order=3, max_splat_count=w.count, field_of_view_y=math.radians(60))
    r = sa.Renderer(dev, cfg)
    scene = sa.Scene(r, w.count)
    scene.load_records(0, w.all())
    f = sa.Frame(dev, 1280, 720)
    cam, view = look_at((0, 0, -5), (0, 0, 0))
    t = []
    for i in range(repeats + 3):
        dev.synchronize()
        t0 = time.perf_counter()
        r.render_frame(f, cam, view, scene)
        dev.synchronize()
        if i >= 3:
            t.append(time.perf_counter() - t0)
    for o in (f, scene, r):
        o.close()
    return float(np.median(t)) * 1e3


SERVING_SCENES = [300_000, 600_000, 1_000_000, 1_500_000, 2_000_000, 3_000_000, 4_500_000, 6_000_000]
SERVING_SIZES = [(640, 480), (1280, 720), (1920, 1080), (3840, 2160)]


def serving_setup(dev, sa, budget_gib: int = 96, in_flight: int = 3, batch: int = 8):
    """A server over 8 resident clustered scenes (0.3M..6M splats)."""
    from splatter_amd.synth import Workload
    cfg = sa.Configuration(spherical_harmonics_order=3, max_splat_count=max(SERVING_SCENES), max_views=batch,
                           tile_pairs_per_splat=8)
    server = sa.Server(dev, cfg, max_in_flight=in_flight, max_width=3840, max_height=2160,
                       memory_budget_bytes=budget_gib << 30)
    t0 = time.perf_counter()
    scenes = []
    for i, n in enumerate(SERVING_SCENES):
        sc = server.create_scene(n)
        for first, rec in Workload("clustered", n, seed=100 + i).chunks():
            sc.load_records(first, rec)
        scenes.append(sc)
    dev.synchronize()
    log(f"serving: {len(scenes)} scenes ({sum(SERVING_SCENES) / 1e6:.1f}M splats) resident in {time.perf_counter() - t0:.1f}s")
    return server, scenes, {"memory_budget_gib": budget_gib, "max_in_flight": in_flight, "batch_width": batch}


def serving(server, scenes, setup: dict, rate: float, seconds: float, warmup_s: float = 1.5, seed: int = 7) -> dict:
    """Poisson stream of render requests over the resident scenes: uniform
    scene, random pose on the radius-6 sphere, resolution uniform over
    640x480..3840x2160.  Latency is the server's submit -> seen-done time of
    every request that arrived in the measured window; throughput is the
    frames finished in that window per second."""
    from splatter_amd.synth import sphere_pose
    rng = np.random.default_rng(seed)
    poses = [sphere_pose(10_000 + i) for i in range(512)]
    lat, batches = [], []
    done_in_window = 0
    failed = 0
    total = warmup_s + seconds
    start = time.perf_counter()
    next_t = rng.exponential(1.0 / rate)
    arrivals = {}
    while True:
        now = time.perf_counter() - start
        while next_t <= now and next_t < total:
            s = int(rng.integers(len(scenes)))
            w, h = SERVING_SIZES[int(rng.integers(len(SERVING_SIZES)))]
            cam, view = poses[int(rng.integers(len(poses)))]
            t = server.submit(scenes[s], w, h, cam, view)
            arrivals[t] = next_t
            next_t += rng.exponential(1.0 / rate)
        for t in server.poll():
            info = server.info(t)
            a = arrivals.pop(t)
            if info["status"] != 0:
                failed += 1
            elif a >= warmup_s:
                lat.append(info["done_ms"] - info["submit_ms"])
                batches.append(info["batch_views"])
            if warmup_s <= (time.perf_counter() - start) <= total:
                done_in_window += 1
            server.release(t)
        if now >= total and not arrivals:
            break
        if now > total + 30:
            raise RuntimeError("serving: requests still unfinished 30 s after the last arrival")
    st = server.stats()
    lat = np.array(lat)
    return {"rate_per_s": rate, "seconds": seconds, "requests": int(len(lat)), "failed": failed,
            "frames_per_s": round(done_in_window / seconds, 2),
            "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
            "mean_batch_views": round(float(np.mean(batches)), 2),
            "memory_peak_gib": round(st["memory_peak"] / 2 ** 30, 2), "slots": st["slots"], **setup,
            "scenes_m": [n / 1e6 for n in SERVING_SCENES], "sizes": [f"{w}x{h}" for w, h in SERVING_SIZES]}


def copy_calibration(nbytes: int = 256 << 20, repeats: int = 3) -> None:
    """Device-to-device copies of a known size: under rocprofv3 --pmc their
    FETCH_SIZE / WRITE_SIZE per dispatch against nbytes give the counters'
    scale on this device."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed for the copy calibration")
    hip.hipMemset(a, 1, ctypes.c_size_t(nbytes))
    for _ in range(repeats):
        hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"copy calibration: {repeats} x {nbytes >> 20} MiB device-to-device")


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=48)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--splats", type=int, default=6_000_000)
    ap.add_argument("--views", type=int, default=16)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--radix-bits", type=int, default=0, help="force the digit width (skips the autotuner)")
    ap.add_argument("--sort-tile", type=int, default=0, help="force keys per sort tile (skips the autotuner)")
    ap.add_argument("--no-autotune", action="store_true", help="8-bit digits, 4096-key sort tiles")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-sort-stress", action="store_true",
                    help="skip the depth sort on the 32M tie-heavy 4K 8-view stress workload")
    ap.add_argument("--in-flight", type=int, default=2,
                    help="batches in flight: the batched loop alternates this many renderers, each with its own "
                         "frames, so one batch's raster overlaps the next one's sort and projection")
    ap.add_argument("--serving-rate", type=float, default=200.0, help="Poisson arrivals per second")
    ap.add_argument("--serving-rate-high", type=float, default=1800.0,
                    help="a second Poisson rate, high enough that requests queue and batches form (0 skips it)")
    ap.add_argument("--serving-seconds", type=float, default=8.0, help="0 skips the serving measurement")
    ap.add_argument("--serving-sweep", type=str, default="",
                    help="comma-separated request rates: run only the serving measurement at each and print them")
    ap.add_argument("--tune-only", action="store_true",
                    help="run the autotuner on the workload and print its result as one JSON line")
    ap.add_argument("--profile", action="store_true",
                    help="batched steps and the stage-timing pass only (every launch has the batch's shape), "
                         "then three 256 MiB device-to-device copies whose counters scale FETCH_SIZE/WRITE_SIZE")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(views and requests shard across GPUs without exchange; run one process per GPU)",
              file=sys.stderr)
        return 2

    import splatter_amd as sa
    from splatter_amd import autotune as at
    from splatter_amd.synth import Workload, load_scene, orbit_poses

    W, H, V_PER = args.width, args.height, args.views
    tuned = None
    if not (args.no_autotune or args.radix_bits or args.sort_tile or args.profile or args.tune_only
            or args.serving_sweep):
        # The tuner creates and frees two renderers per candidate, and with
        # them their streams.  Batches in flight overlapped fully only on
        # renderers whose streams came first in the process: 1540 frames/s
        # against about 1390 after the tuner had run in the same process
        # (MI355X, two batches in flight).  So the tuner runs in a child
        # process, started before this one makes any HIP call.
        cmd = [sys.executable, os.path.abspath(__file__), "--tune-only", "--splats", str(args.splats),
               "--views", str(V_PER), "--width", str(W), "--height", str(H), "--in-flight", str(args.in_flight)]
        child = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=sys.stderr, text=True, timeout=900)
        lines = [x for x in child.stdout.splitlines() if x.startswith("{")]
        if child.returncode != 0 or not lines:
            print(f"autotune child failed with status {child.returncode}", file=sys.stderr)
            return 1
        tuned = at.TuneResult.from_json(json.loads(lines[-1]))
        log(f"autotune picked {tuned.best.key()}")

    if args.serving_sweep:
        dev = sa.Device(0)
        server, scenes, setup = serving_setup(dev, sa)
        for rate in [float(x) for x in args.serving_sweep.split(",")]:
            res = serving(server, scenes, setup, rate, args.serving_seconds)
            log(f"serving at {rate:g}/s: p50 {res['p50_ms']} ms, p99 {res['p99_ms']} ms, "
                f"{res['frames_per_s']} frames/s, mean batch {res['mean_batch_views']}")
            print(json.dumps(res), flush=True)
        return 0

    L = 3
    dev = sa.Device(0)
    cfg = sa.Configuration(spherical_harmonics_order=L, max_splat_count=args.splats, max_views=V_PER,
                           radix_bits_per_digit=args.radix_bits or 8, sort_tile_keys=args.sort_tile or 4096)
    if tuned:
        tuned.apply(cfg)
    in_flight = 1 if (args.profile or args.tune_only) else max(1, args.in_flight)
    if in_flight == 1:
        # renderer 0, on the device stream, takes the batched steps too
        r = sa.Renderer(dev, cfg)
        pipe = [r]
    else:
        # each batch in flight on a renderer with a stream of its own, made
        # before any other renderer (batches on the device stream did not
        # overlap with those of an own-stream renderer: 1282 against 1535
        # frames/s for two in flight); renderer 0 comes after the batched
        # steps
        r = None
        pipe = [sa.Renderer(dev, cfg, own_stream=True) for _ in range(in_flight)]
    wl = Workload("clustered", args.splats, seed=2)
    t0 = time.perf_counter()
    scene = sa.Scene(pipe[0], args.splats)
    positions = np.empty((args.splats, 3), np.float32)
    for first, rec in wl.chunks():
        scene.load_records(first, rec)
        positions[first:first + len(rec)] = rec["center"]
    log(f"scene: {args.splats} splats generated+uploaded in {time.perf_counter() - t0:.1f}s")
    poses = orbit_poses(256)
    if args.tune_only:
        # radix digit width x sort tile (block shape) at this batch width,
        # each candidate timed on the same resident scene over the whole
        # orbit, as the timed steps are (the poses differ in cost, so part
        # of the orbit is a different workload)
        pipe[0].close()
        res = at.autotune_staged(sa, dev, cfg, lambda _r: scene, poses, W, H, V_PER,
                                 steps=max(1, len(poses) // V_PER), log=log,
                                 in_flight=max(1, args.in_flight))
        print(json.dumps(res.to_json()), flush=True)
        return 0
    frames = [sa.Frame(dev, W, H) for _ in range(V_PER)]
    pipe_frames = [frames] + [[sa.Frame(dev, W, H) for _ in range(V_PER)] for _ in range(len(pipe) - 1)]

    def sync_all():
        dev.synchronize()
        for rr in pipe:
            if rr.own_stream:
                rr.synchronize()

    def batch(k, renderers=None):
        renderers = renderers or pipe
        j = k % len(renderers)
        sel = [poses[(k * V_PER + i) % len(poses)] for i in range(V_PER)]
        renderers[j].render_views(pipe_frames[j], [p[0] for p in sel], [p[1] for p in sel], scene)

    for k in range(max(args.warmup, in_flight)):
        batch(k)
    sync_all()

    # batched: K steps of V_PER views, one graph replay per step, in_flight
    # steps on the device at a time
    t0 = time.perf_counter()
    for k in range(args.steps):
        batch(k)
    sync_all()
    t_batched = time.perf_counter() - t0
    batched_fps = args.steps * V_PER / t_batched
    log(f"batched: {batched_fps:.1f} frames/s")

    # per-stage device time per frame: HIP events around every stage of every
    # batch of a second timed region of the same work on the first batched
    # renderer alone (direct launches, so each stage's events bracket it alone)
    t_r = pipe[0]
    t_r.enable_timing(True)
    for k in range(args.steps):
        batch(k, [t_r])
    stage_ms = t_r.stage_times()
    t_r.enable_timing(False)

    if args.profile:
        copy_calibration()
        print(json.dumps({"profile": True, "batched_frames_per_s": round(batched_fps, 2), "views_per_batch": V_PER,
                          "stage_ms_per_frame": {s: round(float(stage_ms[i]), 4) for i, s in enumerate(STAGES)}}))
        return 0

    # renderer 0 on the device stream: the plain path, the stage timing and
    # the counters
    if r is None:
        r = sa.Renderer(dev, cfg)
    # plain: the same views, one render_frame each
    n_plain = args.steps * V_PER
    for i in range(2):
        r.render_frame(frames[0], *poses[i], scene)
    dev.synchronize()
    t0 = time.perf_counter()
    for i in range(n_plain):
        cam, view = poses[i % len(poses)]
        r.render_frame(frames[i % V_PER], cam, view, scene)
    dev.synchronize()
    plain_fps = n_plain / (time.perf_counter() - t0)
    log(f"plain: {plain_fps:.1f} frames/s")

    for k, rr in enumerate(pipe):
        if rr is not r:
            rr.close()
            if k > 0:
                for f in pipe_frames[k]:
                    f.close()
    pipe, pipe_frames = [r], [frames]

    r.enable_timing(True)
    for i in range(min(n_plain, 64)):
        r.render_frame(frames[0], *poses[i % len(poses)], scene)
    plain_stage_ms = r.stage_times()
    r.enable_timing(False)

    # latency of single frames (submit -> frame ready), and V / K per view
    lat, Vs, Ks = [], [], []
    for i in range(min(64, len(poses))):
        cam, view = poses[(i * 4) % len(poses)]
        dev.synchronize()
        t1 = time.perf_counter()
        r.render_frame(frames[0], cam, view, scene)
        dev.synchronize()
        lat.append(time.perf_counter() - t1)
        Vs.append(r.visible_count())
        Ks.append(r.pair_count())
    V, K = float(np.mean(Vs)), float(np.mean(Ks))
    r.enable_tile_stats(True)
    r.render_frame(frames[0], *poses[0], scene)
    ts = r.tile_stats()
    r.enable_tile_stats(False)
    tiles_info = {"max_list": int(ts[:, 0].max()), "mean_list": round(float(ts[:, 0].mean()), 1),
                  "max_consumed": int(ts[:, 1].max()), "consumed_total": int(ts[:, 1].sum()),
                  "pairs_total": int(ts[:, 0].sum())}
    tiles = math.ceil(W / 16) * math.ceil(H / 16)
    P = math.ceil(30 / cfg.radix_bits_per_digit)
    Q = 1 if tiles <= 256 else (2 if tiles <= 65536 else 3)
    nbytes = stage_bytes(V, K, args.splats, W, H, L, P, Q, V_PER)
    per_stage = {s: {"ms": round(float(stage_ms[i]), 4), "plain_ms": round(float(plain_stage_ms[i]), 4),
                     "GB/s": round(nbytes[s] / (stage_ms[i] * 1e-3) / 1e9, 1) if stage_ms[i] > 0 else None}
                 for i, s in enumerate(STAGES)}
    dom = max(range(len(STAGES)), key=lambda i: stage_ms[i])
    dom_name = STAGES[dom]
    launches = {"depth_sort": P, "pair_sort": Q}.get(dom_name, 1)
    achieved = nbytes[dom_name] / (stage_ms[dom] * 1e-3) / 1e9
    # the raster is bound by VALU issue rather than HBM: its wave instructions
    # per launch (counters of the same workload) over the live launch time
    valu = measured_valu(dom_name)
    valu_issue = None
    if valu and launches == 1:
        rate = valu[1] / (stage_ms[dom] * 1e-3 * V_PER)
        valu_issue = {"kernel": valu[0], "instructions_per_launch": valu[1], "achieved": round(rate / 1e9, 1),
                      "peak": round(VALU_PEAK_PER_S / 1e9, 1), "unit": "G wave64 VALU instructions/s",
                      "frac": round(rate / VALU_PEAK_PER_S, 4),
                      "source": "SQ_INSTS_VALU per launch in profiles/kernels_6M_clustered_1080p_16views.md"}
    sort_i = STAGES.index("depth_sort")
    sort_gkeys = V / (stage_ms[sort_i] * 1e-3) / 1e9 if stage_ms[sort_i] > 0 else None

    serve = serve_high = None
    if args.serving_seconds > 0:
        r.close()
        scene.close()
        server, scenes, setup = serving_setup(dev, sa)
        serve = serving(server, scenes, setup, args.serving_rate, args.serving_seconds)
        log(f"serving at {args.serving_rate}/s: p50 {serve['p50_ms']} ms, p99 {serve['p99_ms']} ms, "
            f"{serve['frames_per_s']} frames/s")
        if args.serving_rate_high > 0:
            serve_high = serving(server, scenes, setup, args.serving_rate_high, args.serving_seconds)
            log(f"serving at {args.serving_rate_high}/s: p50 {serve_high['p50_ms']} ms, p99 {serve_high['p99_ms']} ms, "
                f"{serve_high['frames_per_s']} frames/s, mean batch {serve_high['mean_batch_views']}")
        server.close()

    stress = None
    if not args.no_sort_stress:
        stress = sort_stress(sa, dev, log)

    cpu = None
    if not args.no_cpu:
        c0_gpu = config0_gpu_ms(sa, dev)
        cpu = cpu_baseline(positions, poses[::32], W, H)
        cpu["config0_gpu_frame_ms"] = round(c0_gpu, 3)
        log(f"cpu baseline: {cpu['value']} frames/s on 1 core, {cpu['value_all_cores']} on {cpu['cores_all']}; "
            f"small frame {cpu['config0_frame_ms']} ms on the CPU, {cpu['config0_gpu_frame_ms']} ms on the GPU")

    out = {
        "metric": METRIC,
        "value": round(batched_fps, 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_batched / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{args.splats // 1_000_000}M clustered splats, SH3, {W}x{H}, "
                               f"256-pose orbit, {V_PER} views per batch",
                   "splats": args.splats, "width": W, "height": H, "views_per_batch": V_PER,
                   "radix_bits_per_digit": cfg.radix_bits_per_digit, "sort_tile_keys": cfg.sort_tile_keys,
                   "raster_batch": cfg.raster_batch, "raster_longest_first": cfg.raster_longest_first,
                   "raster_mode": cfg.raster_mode, "batches_in_flight": in_flight,
                   "autotune": tuned.to_json() if tuned else None},
        "plain_frames_per_s": round(plain_fps, 2),
        "batched_over_plain": round(batched_fps / plain_fps, 3),
        "sort_gkeys_per_s": round(sort_gkeys, 2) if sort_gkeys else None,
        "latency_ms": ({"p50": serve["p50_ms"], "p99": serve["p99_ms"],
                        "kind": f"serving, Poisson {serve['rate_per_s']:g} requests/s, submit to done"}
                       if serve else None),
        "plain_frame_latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 3),
                                   "p99": round(float(np.percentile(lat, 99)) * 1e3, 3),
                                   "kind": "single plain frame, submit to ready, back to back"},
        "serving": serve,
        "serving_high_rate": serve_high,
        "visible_per_view": int(V),
        "pairs_per_view": int(K),
        "stages": per_stage,
        "raster_tiles": tiles_info,
        "sort_stress": stress,
        "roofline": {"bound": "hbm", "kernel": dom_name, "launches_per_batch": launches,
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(dom_name),
                     "valu_issue": valu_issue},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
