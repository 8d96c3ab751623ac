#!/usr/bin/env python3
"""Benchmark: street-scale single-view 3DGS training on one MI355X.

Workload (the headline configuration): 475,115 Gaussians, SH degree 3, fp32, means on
spheres around two capture centres 10 m apart with log-uniform depth in 2-60 m, scales
from the exact 3-NN mean distance, 1264x832 pinhole views (fx = fy = 900) on the 20 m
path between the centres, seeded ground-truth Gaussians rendered as targets; one view
per step, densify/prune every 100 steps from step 500 (grad 2e-4, opacity prune 0.005),
3000-step schedule. A step is one msplat_hip_train_step plus the trainer's schedule.

    python bench.py [--steps K] [--warmup W]

prints one JSON line. The timed region is K steps of the schedule after W warm-up
steps, bracketed by stream syncs; per-stage GPU times come from HIP events on the
training stream inside that region.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "msplat-street_amd"))

import numpy as np  # noqa: E402

METRIC = "train iter/s @475k Gaussians 1264x832; batched views/s; served p99 latency"
N_GAUSSIANS = 475_115
W_IMG, H_IMG, FOCAL = 1264, 832, 900.0
N_VIEWS = 64
BG = (0.6130, 0.0101, 0.3984)
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
# Algorithmic bytes per unit (minimum fp32 reads + writes), per kernel class
BYTES = {
    "rast_fwd": (36, 20, 0),            # per intersection, per pixel, per gaussian
    "rast_bwd": (4, 28, 40),            # + SLOT_BYTES per written gradient slot
    "loss_fwd_bwd": (0, 36, 0),
    "prefix_sort_pack": (24 + 144 + 8, 0, 24),
}
DOMINANT = "rast_bwd"  # the train step's longest stage (the raster backward)
SLOT_BYTES = 37  # a written gradient slot: nine floats and its flag (all-zero slots are only flagged)


def slots_written(ctx) -> int:
    """The context's running count of gradient slots written by the raster backward."""
    return int(ctx.download(ctx.frame().slots_written, (1,), np.uint64)[0])


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_workload(ctx, seed=0):
    from msplat._abi import DeviceArray
    from msplat.backend import DeviceGaussians, render
    from msplat.scenes import street_cameras, street_points
    pts, rgb = street_points(seed, N_GAUSSIANS)
    dg = DeviceGaussians.from_points(ctx, pts, rgb, degree=3, capacity=4 * N_GAUSSIANS)
    # ground truth: a separate seeded set, made opaque and view-dependent
    gpts, grgb = street_points(seed + 1, N_GAUSSIANS)
    gt = DeviceGaussians.from_points(ctx, gpts, grgb, degree=3)
    rng = np.random.default_rng(seed + 2)
    gt.params["opac"].upload(rng.normal(1.5, 1.0, (N_GAUSSIANS, 1)).astype(np.float32))
    gt.params["rest"].upload(rng.normal(0, 0.05, (N_GAUSSIANS, 45)).astype(np.float32))
    cams = street_cameras(N_VIEWS, W_IMG, H_IMG, FOCAL, FOCAL)
    bg = (0.6130, 0.0101, 0.3984)
    targets = []
    for c in cams:
        p = render(ctx, gt, dict(c, dtu=3), bg)
        t = DeviceArray(ctx, (H_IMG, W_IMG, 3))
        ctx.upload(t.ptr, ctx.download(p, (H_IMG, W_IMG, 3), np.float32))
        targets.append(t)
    return dg, gt, cams, targets


BATCH_SIZES = [(640, 360), (1264, 832), (1920, 1080), (2560, 1440)]
BATCH_N = 1_000_000
BATCH_VIEWS = 16
BATCH_POOL = 32


def psnr(a, b):
    mse = float(np.mean((np.asarray(a, np.float64) - np.asarray(b, np.float64)) ** 2))
    return 10.0 * np.log10(1.0 / max(mse, 1e-20))


def bench_batched(steps=8, warmup=2, seed=10):
    """Batched views: 1,000,000 SH3 Gaussians (street recipe), 16 views per launch drawn
    from a seeded pool of 32 (pose on the street path, resolution uniform over 640x360,
    1264x832, 1920x1080, 2560x1440; fx = 900 W / 1264). The batched step (one
    msplat_hip_train_step_batch, replayed from its captured graph; gradients summed over
    the 16 views, one Adam step) against the plain form (16 msplat_hip_train_step, one
    view per launch), and msplat_hip_render_batch against 16 msplat_hip_render, on the
    same views. ΔPSNR compares the batched and plain renders of every pooled view."""
    from msplat._abi import Context, DeviceArray
    from msplat.backend import (BatchViews, DeviceGaussians, adam_coeffs, render, render_batch, render_batch_numpy,
                                render_numpy, train_step, train_step_batch)
    from msplat.camera import make_camera
    from msplat.scenes import street_cameras, street_points
    ctx = Context(0)
    rng = np.random.default_rng(seed)
    t0 = time.perf_counter()
    pts, rgb = street_points(seed, BATCH_N)
    gpts, grgb = street_points(seed + 1, BATCH_N)
    gt_set = DeviceGaussians.from_points(ctx, gpts, grgb, degree=3)
    gt_set.params["opac"].upload(rng.normal(1.5, 1.0, (BATCH_N, 1)).astype(np.float32))
    gt_set.params["rest"].upload(rng.normal(0, 0.05, (BATCH_N, 45)).astype(np.float32))
    pool, targets = [], []
    sizes = [BATCH_SIZES[i] for i in rng.integers(0, len(BATCH_SIZES), BATCH_POOL)]
    path = street_cameras(BATCH_POOL, 64, 64)
    for k, (W, H) in enumerate(sizes):
        f = 900.0 * W / 1264.0
        cam = make_camera(path[k]["cam_to_world"], W, H, f, f, dtu=3)
        pool.append(cam)
        t = DeviceArray(ctx, (H, W, 3))
        ctx.upload(t.ptr, render_numpy(ctx, gt_set, cam, BG))
        targets.append(t)
    del gt_set
    bg = [BG]
    lrs = [1.6e-4, 5e-3, 1e-3, 2.5e-3, 1.25e-4, 5e-2]
    picks = [rng.permutation(BATCH_POOL)[:BATCH_VIEWS] for _ in range(warmup + steps)]
    mpx = [sum(pool[i]["W"] * pool[i]["H"] for i in p) / 1e6 for p in picks]

    def fresh():
        return DeviceGaussians.from_points(ctx, pts, rgb, degree=3)

    # batched training, graph replays
    dg = fresh()
    ctx.set_batch_graph(True)
    for t, p in enumerate(picks[:warmup], start=1):
        s_, b_ = adam_coeffs(lrs, t)
        train_step_batch(ctx, dg, BatchViews([pool[i] for i in p], 3, bg, [targets[i] for i in p]), s_, b_)
    ctx.sync()
    tb = time.perf_counter()
    for t, p in enumerate(picks[warmup:], start=warmup + 1):
        s_, b_ = adam_coeffs(lrs, t)
        train_step_batch(ctx, dg, BatchViews([pool[i] for i in p], 3, bg, [targets[i] for i in p]), s_, b_)
    ctx.sync()
    batched_s = time.perf_counter() - tb
    captures, replays = ctx.batch_graph_stats()
    # plain training, one view per launch
    dg = fresh()
    t = 0
    for p in picks[:warmup]:
        for i in p:
            t += 1
            s_, b_ = adam_coeffs(lrs, t)
            train_step(ctx, dg, pool[i], BG, targets[i], s_, b_, read_loss=False)
    ctx.sync()
    tp = time.perf_counter()
    for p in picks[warmup:]:
        for i in p:
            t += 1
            s_, b_ = adam_coeffs(lrs, t)
            train_step(ctx, dg, pool[i], BG, targets[i], s_, b_, read_loss=False)
    ctx.sync()
    plain_s = time.perf_counter() - tp
    # rendering: batched vs plain on the same views, and PSNR of both against the targets
    bvs = [BatchViews([pool[i] for i in p], 3, bg) for p in picks]
    for b in bvs[:warmup]:
        render_batch(ctx, dg, b)
    ctx.sync()
    tr0 = time.perf_counter()
    for b in bvs[warmup:]:
        render_batch(ctx, dg, b)
    ctx.sync()
    rb_s = time.perf_counter() - tr0
    tr1 = time.perf_counter()
    for p in picks[warmup:]:
        for i in p:
            render(ctx, dg, pool[i], BG)
    ctx.sync()
    rp_s = time.perf_counter() - tr1
    dpsnr = 0.0
    whole = BatchViews(pool[:BATCH_VIEWS], 3, bg)
    for img, cam, tgt in zip(render_batch_numpy(ctx, dg, whole), pool[:BATCH_VIEWS], targets[:BATCH_VIEWS]):
        ref = tgt.numpy()
        dpsnr = max(dpsnr, abs(psnr(img, ref) - psnr(render_numpy(ctx, dg, cam, BG), ref)))
    views = BATCH_VIEWS * steps
    out = {
        "workload": "1,000,000 SH3 Gaussians, 16 views per launch from a pool of 32 (resolutions uniform over "
                    "640x360/1264x832/1920x1080/2560x1440)",
        "mean_mpx_per_launch": round(float(np.mean(mpx[warmup:])), 2),
        "train_views_per_s": {"batched": round(views / batched_s, 2), "plain": round(views / plain_s, 2),
                              "ratio": round(plain_s / batched_s, 3)},
        "render_views_per_s": {"batched": round(views / rb_s, 2), "plain": round(views / rp_s, 2),
                               "ratio": round(rp_s / rb_s, 3)},
        "delta_psnr_db": round(dpsnr, 4),
        "graph": {"captures": captures, "replays": replays},
        "launches_timed": steps, "setup_s": round(tb - t0, 1),
    }
    ctx.close()
    return out


SERVE_SCENES = (500_000, 1_000_000, 2_000_000, 3_000_000)
SERVE_RATES = (200, 500, 1000, 2000)


def cluster_cloud(seed, n, center, clusters=64, sigma=0.5, box=8.0):
    """A Gaussian mixture of `clusters` clusters (sigma in m) with centres uniform in a
    box around `center`, and a colour per cluster."""
    rng = np.random.default_rng(seed)
    cc = np.asarray(center) + rng.uniform(-box / 2, box / 2, (clusters, 3))
    col = rng.integers(30, 226, (clusters, 3))
    lab = rng.integers(0, clusters, n)
    pts = (cc[lab] + rng.normal(0, sigma, (n, 3))).astype(np.float32)
    rgb = np.clip(col[lab] + rng.integers(-25, 26, (n, 3)), 0, 255).astype(np.uint8)
    return pts, rgb


def bench_serving(rates=SERVE_RATES, duration=1.0, slots=2, max_views=16, seed=20):
    """Serving: four resident SH3 scenes (0.5M, 1M, 2M, 3M Gaussians, 64-cluster clouds in
    an 8 m box, kNN scales, opacity logits N(1.5, 1), higher SH bands N(0, 0.05)).
    Poisson arrivals at each rate for `duration` s; pose on a sphere of radius 3-10 m
    around the scene's centre, resolution 640x360 40% / 1280x720 30% / 1920x1080 20% /
    3840x2160 10%. The continuous-batching server (2 slots, up to 16 views per batch,
    HBM-budget admission) against the plain form (one msplat_hip_render per request,
    first come first served), each followed by the RGBA8 pack."""
    from msplat._abi import Context
    from msplat.backend import DeviceGaussians
    from msplat.serving import Server, latency_stats, poisson_requests, serve_plain
    ctx = Context(0)
    t0 = time.perf_counter()
    rng = np.random.default_rng(seed)
    centers = [(60.0 * k, 0.0, 0.0) for k in range(len(SERVE_SCENES))]
    scenes = []
    for k, (n, c) in enumerate(zip(SERVE_SCENES, centers)):
        pts, rgb = cluster_cloud(seed + k, n, c)
        dg = DeviceGaussians.from_points(ctx, pts, rgb, degree=3)
        dg.params["opac"].upload(rng.normal(1.5, 1.0, (n, 1)).astype(np.float32))
        dg.params["rest"].upload(rng.normal(0, 0.05, (n, 45)).astype(np.float32))
        scenes.append(dg)
    ctx.sync()
    setup = time.perf_counter() - t0
    srv = Server(scenes, slots=slots, max_views=max_views, hbm_budget=200e9, background=BG)
    warm = poisson_requests(2000.0, 0.1, centers, seed=seed + 99)
    srv.run(warm)
    serve_plain(scenes, poisson_requests(500.0, 0.1, centers, seed=seed + 98), background=BG)
    out = {"workload": "4 resident SH3 scenes (0.5M/1M/2M/3M Gaussians); Poisson arrivals; 640x360 40% / "
                       "1280x720 30% / 1920x1080 20% / 3840x2160 10%; pose radius 3-10 m",
           "slots": slots, "max_views": max_views, "duration_s": duration, "setup_s": round(setup, 1), "rates": {}}
    for rate in rates:
        reqs = poisson_requests(float(rate), duration, centers, seed=seed + rate)
        b = latency_stats(srv.run(reqs))
        b["rejected"] = len(srv.rejected)
        plain = latency_stats(serve_plain(scenes, poisson_requests(float(rate), duration, centers, seed=seed + rate),
                                          background=BG))
        key = "1920x1080"
        out["rates"][str(rate)] = {
            "requests": len(reqs),
            "batched": b, "plain": plain,
            "p99_1080p_ms": {"batched": b["by_resolution"].get(key, {}).get("p99_ms"),
                             "plain": plain["by_resolution"].get(key, {}).get("p99_ms")},
        }
        log(f"serving {rate} req/s: batched {b['views_per_s']} views/s p50 {b['p50_ms']} p99 {b['p99_ms']} ms; "
            f"plain {plain['views_per_s']} views/s p50 {plain['p50_ms']} p99 {plain['p99_ms']} ms")
    srv.close()
    ctx.close()
    return out


def tune_train_step(ctx, cams, targets, seed=5):
    """Picks the rasterizer variants (msplat.autotune) on the headline train step, run on
    a scratch copy of the scene (the same street recipe) so the trainer's schedule is not
    advanced; the choice stays set on the context."""
    from msplat import autotune
    from msplat.backend import DeviceGaussians, adam_coeffs, train_step
    from msplat.scenes import street_points
    pts, rgb = street_points(seed, N_GAUSSIANS)
    scratch = DeviceGaussians.from_points(ctx, pts, rgb, degree=3)
    lrs = [1.6e-4, 5e-3, 1e-3, 2.5e-3, 1.25e-4, 5e-2]
    k = [0]

    def step():
        k[0] += 1
        s_, b_ = adam_coeffs(lrs, k[0])
        i = k[0] % len(cams)
        train_step(ctx, scratch, dict(cams[i], dtu=3), BG, targets[i], s_, b_, read_loss=False)

    res = autotune.autotune(ctx, step, warmup=3, repeats=12)
    del scratch
    return res


def copy_calibration(mib=256, times=5):
    """Copies `mib` MiB device to device `times` times through msplat_hip_memcpy: a kernel
    of known traffic (mib MiB read + mib MiB written per dispatch) for scaling the HBM
    counters of a profile."""
    import ctypes as C
    from msplat._abi import Context, DeviceArray, check, lib
    ctx = Context(0)
    n = mib * 1024 * 1024
    a, b = DeviceArray(ctx, (n,), np.uint8), DeviceArray(ctx, (n,), np.uint8)
    for _ in range(times):
        check(lib().msplat_hip_memcpy(ctx.ptr, C.c_void_p(b.ptr), C.c_void_p(a.ptr), n, 2))
    ctx.sync()
    log(f"copied {times} x {mib} MiB")


def profiled_traffic(kernel_stage, tuning):
    """HBM bytes per launch of the dominant kernel, from the committed counter summary
    (profiles/raster_backward_street_475k_traffic.json, FETCH_SIZE and WRITE_SIZE scaled
    by a copy of known size, per kernel variant) for the timed variant; else None."""
    path = os.path.join(ROOT, "profiles", "raster_backward_street_475k_traffic.json")
    if kernel_stage != "rast_bwd" or not os.path.exists(path):
        return None
    prof = json.load(open(path))
    variant = "raster_backward_kernel<%d, %d>" % (tuning.get("bwd_ppt", 4), tuning.get("bwd_batch", 128))
    entry = prof.get("variants", {}).get(variant)
    return entry["bytes_per_launch"] if entry else None


def cpu_baseline_sample(dg, cams, targets, ctx, step_index, seconds=20.0):
    """Times the reference CPU train step (oracle/_ref, perf build) — or the oracle when
    the reference was not built — on this host, on the workload's own scene and view."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    threads = int(os.environ.get("MSPLAT_THREADS", "16"))
    os.environ["MSPLAT_THREADS"] = str(threads)
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    from oracle import Oracle, Reference, reference_available
    kind = "reference" if reference_available(perf=True) else "port"
    impl = Reference(perf=True) if kind == "reference" else Oracle()
    full = dg.to_numpy()
    keys = ("means", "scales", "quats", "dc", "rest", "opac")
    st = {k: full[k].copy() for k in keys}
    for k in keys:
        st["m1_" + k] = np.zeros_like(full[k])
        st["m2_" + k] = np.zeros_like(full[k])
    n = full["means"].shape[0]
    st.update(degree=dg.degree, vis=np.zeros(n, np.float32), gn=np.zeros(n, np.float32), ms=np.zeros(n, np.float32))
    cam = dict(cams[0], dtu=min(step_index // 1000, 3))
    gt = targets[0].numpy()
    lrs = np.array([1.6e-4, 5e-3, 1e-3, 2.5e-3, 1.25e-4, 5e-2], np.float32)
    steps, bc2 = lrs / np.float32(0.1), np.full(6, np.sqrt(np.float32(0.001)), np.float32)
    times = []
    t_start = time.perf_counter()
    while True:
        t0 = time.perf_counter()
        impl.train_step(st, cam, np.array([0.6130, 0.0101, 0.3984], np.float32), gt, steps, bc2)
        times.append(time.perf_counter() - t0)
        if time.perf_counter() - t_start > seconds or len(times) >= 5:
            break
    per = float(np.median(times))
    return {"value": 1.0 / per, "unit": "iter/s", "cores": threads, "kind": kind,
            "sample": f"{len(times)} train step(s) of the same 475,115-Gaussian scene on view 0 at "
                      f"{W_IMG}x{H_IMG} (SH degree in use {cam['dtu']}), median {per:.2f} s/step, "
                      f"MSPLAT_THREADS={threads}, host {platform.processor() or platform.machine()}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2950)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--iterations", type=int, default=3000, help="schedule length (densify window, LR decay)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--bwd-ppt", type=int, default=0, help="backward pixels per thread (0: library default)")
    ap.add_argument("--bwd-batch", type=int, default=0, help="backward LDS batch depth (0: library default)")
    ap.add_argument("--fwd-batch", type=int, default=0, help="forward LDS batch depth (0: library default)")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-views comparison")
    ap.add_argument("--batched-steps", type=int, default=8)
    ap.add_argument("--no-serving", action="store_true", help="skip the serving sweep")
    ap.add_argument("--no-autotune", action="store_true", help="keep the library's default kernel variants")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 256 MiB device to device five times (the known-size kernel that scales the "
                         "FETCH_SIZE / WRITE_SIZE counters) and exit")
    ap.add_argument("--serving-seconds", type=float, default=1.0)
    ap.add_argument("--stage-steps", type=int, default=200,
                    help="steps after the timed ones that time every stage for the breakdown")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus})",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return

    from msplat._abi import Context
    from msplat.trainer import TrainConfig, Trainer

    ctx = Context(0)
    pinned = args.bwd_ppt or args.bwd_batch or args.fwd_batch  # a fixed variant: no autotuning
    if pinned:
        from msplat._abi import TUNE_BWD_BATCH, TUNE_BWD_PPT, TUNE_FWD_BATCH
        for key, value in ((TUNE_BWD_PPT, args.bwd_ppt), (TUNE_BWD_BATCH, args.bwd_batch),
                           (TUNE_FWD_BATCH, args.fwd_batch)):
            if value:
                ctx.set_tuning(key, value)
    t_setup = time.perf_counter()
    dg, gt, cams, targets = build_workload(ctx)
    del gt
    ctx.sync()
    setup_s = time.perf_counter() - t_setup
    log(f"workload ready in {setup_s:.1f} s: {dg.n} gaussians, {len(cams)} views")
    tuning = None
    if not args.no_autotune and not pinned:
        tuning = tune_train_step(ctx, cams, targets)
        log("autotuned: " + str({k: tuning[k] for k in ("bwd_ppt", "fwd_batch", "bwd_batch", "ms")}))
    cfg = TrainConfig(iterations=args.iterations, sh_degree=3, warmup_length=499, cull_alpha_thresh=0.005)
    tr = Trainer(ctx, dg, cams, targets, cfg)
    n_start = dg.n
    for _ in range(args.warmup):
        tr.step()
    ctx.sync()
    first_timed_step = tr.step_count + 1
    log(f"warm-up done ({args.warmup} steps); timing {args.steps} steps")

    # The timed steps record events around the dominant stage only: an event between two
    # kernels costs the stream a few microseconds, and six stages' worth would take
    # several percent off the step rate.
    ctx.enable_stage_timing(True)
    ctx.set_stage_mask((DOMINANT,))
    ctx.drain_stage_times()
    m_total, n_total = 0, 0
    ctx.sync()
    written0 = slots_written(ctx)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        tr.step()
        m_total += ctx.frame().num_isects
        n_total += dg.n
    ctx.sync()
    elapsed = time.perf_counter() - t0
    written = slots_written(ctx) - written0
    timed = ctx.drain_stage_times()
    # stage breakdown over a separate window of steps that times every stage
    ctx.set_stage_mask(None)
    for _ in range(args.stage_steps):
        tr.step()
    stages = ctx.drain_stage_times()
    ctx.enable_stage_timing(False)
    loss = tr.step(read_loss=True)
    log(f"timed {args.steps} steps in {elapsed:.3f} s; gaussians {n_start} -> {dg.n}")

    K = args.steps
    iters = K / elapsed
    P = W_IMG * H_IMG
    # dominant kernel class and its roofline, from the events of the timed steps
    dom = DOMINANT
    heaviest = max(stages, key=lambda k: stages[k][0])
    if heaviest != dom:
        log(f"note: the stage breakdown's heaviest stage is {heaviest}, not {dom}")
    dom_ms, dom_calls = timed[dom]
    avg_ms = dom_ms / max(dom_calls, 1)
    per_isect, per_px, per_g = BYTES.get(dom, (0, 0, 0))
    per_slot = SLOT_BYTES if dom == "rast_bwd" else 0
    alg_bytes = (per_isect * m_total + per_px * P * K + per_g * n_total + per_slot * written) / max(dom_calls, 1)
    achieved = alg_bytes / (avg_ms * 1e-3) / 1e9
    from msplat import autotune
    cur = autotune.current(ctx)
    traffic = profiled_traffic(dom, cur)
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "traffic_source": "profiles/raster_backward_street_475k_traffic.md" if traffic else None,
                "note": ("the raster backward is bound by its per-pixel VALU work, not HBM: its VALU instructions "
                         "take 82% of its time (profiles/raster_backward_street_475k_traffic.md); all-zero gradient "
                         "slots are flagged, not written, so its algorithmic bytes count written slots only")
                if dom == "rast_bwd" else None,
                "variant": cur,
                "avg_launch_ms": round(avg_ms, 4), "algorithmic_bytes_per_launch": int(alg_bytes),
                "intersections_per_launch": int(m_total / max(dom_calls, 1)),
                "slots_written_per_launch": int(written / max(dom_calls, 1)) if per_slot else None,
                "stages_ms_per_step": {k: round(v[0] / max(v[1], 1), 4) for k, v in stages.items()},
                "stages_window_steps": args.stage_steps}
    batched = None
    if not args.no_batched:
        log("batched views: 1M gaussians, 16 views per launch")
        batched = bench_batched(args.batched_steps)
        log(f"batched: {batched['train_views_per_s']} train views/s, {batched['render_views_per_s']} render views/s")
    serving = None
    if not args.no_serving:
        log("serving: 4 resident scenes, Poisson arrivals")
        serving = bench_serving(duration=args.serving_seconds)
    cpu = None
    if not args.no_cpu_baseline:
        log("timing the CPU baseline")
        cpu = cpu_baseline_sample(dg, cams, targets, ctx, first_timed_step, args.cpu_seconds)
    out = {
        "metric": METRIC, "value": round(iters, 3), "unit": "iter/s", "n_gpus": 1, "steps": K,
        "warmup": args.warmup, "ms_per_step": round(1e3 * elapsed / K, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "street-scale single-view training: 475,115 SH3 Gaussians, 1264x832, one view "
                               "per step, densify every 100 from step 500 (grad 2e-4, prune 0.005), 3000-step "
                               "schedule",
                   "gaussians_start": n_start, "gaussians_end": dg.n, "steps_timed": [first_timed_step,
                                                                                      first_timed_step + K - 1],
                   "mean_intersections_per_step": int(m_total / K), "densify_events": tr.densify_log,
                   "final_loss": loss, "setup_s": round(setup_s, 1),
                   "tuning": {k: tuning[k] for k in ("bwd_ppt", "fwd_batch", "bwd_batch")} if tuning else None},
        "roofline": roofline,
        "batched": batched,
        "serving": serving,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
