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
FP32_PEAK_TFLOPS = 157.3   # published FP32 vector peak (it counts packed dual FMAs)
SIMDS, CLOCK_HZ = 1024, 2.4e9  # 256 CUs x 4 SIMDs; a wave64 VALU instruction issues over 4 cycles
DOMINANT = "rast_bwd"  # the train step's longest stage (the raster backward)
SLOT_BYTES = 37  # a written gradient slot: nine floats and its flag (all-zero slots are only flagged)
COUNTERS = os.path.join(ROOT, "profiles", "train_step_street_475k_counters.json")


def alg_bytes_8d(stage, M, V, N, P, T, B):
    """Minimal fp32 bytes per step of a stage, by the per-unit formulas of the path's
    measurement plan (M intersections, V visible and N Gaussians, P pixels, T tiles, B SH
    bases in use):
      proj_sh_fwd       52 + 12 (B - 1) read + 48 written per Gaussian;
      prefix_sort_pack  key generation 24 per visible Gaussian + 12 per intersection, a
                        six-pass 64-bit LSD sort 144, ranges 8 per intersection + 8 per
                        tile, packing 72 per intersection;
      rast_fwd          36 per intersection + 20 per pixel;
      loss_fwd_bwd      36 per pixel;
      rast_bwd          40 per intersection + 28 per pixel + 40 per Gaussian;
      proj_sh_bwd_adam  104 + 72 B per visible Gaussian, dense Adam 11 x 28 and gradient
                        statistics 36 per Gaussian."""
    return {
        "proj_sh_fwd": (52 + 12 * (B - 1) + 48) * N,
        "prefix_sort_pack": 24 * V + (12 + 144 + 8 + 72) * M + 8 * T,
        "rast_fwd": 36 * M + 20 * P,
        "loss_fwd_bwd": 36 * P,
        "rast_bwd": 40 * M + 28 * P + 40 * N,
        "proj_sh_bwd_adam": (104 + 72 * B) * V + (11 * 28 + 36) * N,
    }[stage]


def alg_bytes_build(stage, M, V, N, P, T, B, written=0):
    """The same minimum for the way this build does the stage, where it differs: the sort
    stage sorts V depth keys (4 passes x 24 B), emits one 4-byte key and a flag byte per
    intersection, sorts them by tile in 2 passes (12 B each) and finds ranges (4 B per
    intersection + 4 per tile), without packing; the raster backward reads 4 B per
    intersection and writes 37 B per written gradient slot (an all-zero slot is flagged,
    not written), plus 28 per pixel and 40 per Gaussian."""
    if stage == "prefix_sort_pack":
        return 96 * V + (5 + 24 + 4) * M + 4 * T
    if stage == "rast_bwd":
        return 4 * M + SLOT_BYTES * written + 28 * P + 40 * N
    return alg_bytes_8d(stage, M, V, N, P, T, B)


def slots_written(ctx) -> int:
    """The context's running count of gradient slots written by the raster backward."""
    return int(ctx.download(ctx.frame().slots_written, (1,), np.uint64)[0])


def frame_counts(ctx) -> tuple:
    """The context's running counts of one-view intersections and visible gaussians."""
    f = ctx.frame()
    return (int(ctx.download(f.isects_running, (1,), np.uint64)[0]),
            int(ctx.download(f.visible_running, (1,), np.uint64)[0]))


def time_step_graph(ctx, tr, windows=4, steps=50):
    """Alternating windows of launch-by-launch steps and replays of the captured step
    (msplat_hip_set_step_graph) on the same schedule; each mode's first step (which sizes
    the capacity) is outside its window. Returns both rates."""
    rates = {False: [0.0, 0], True: [0.0, 0]}
    c0 = ctx.step_graph_stats()
    for w in range(2 * windows):
        graph = w % 2 == 1
        ctx.set_step_graph(graph)
        tr.step()
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            tr.step()
        ctx.sync()
        rates[graph][0] += time.perf_counter() - t0
        rates[graph][1] += steps
    ctx.set_step_graph(False)
    c1 = ctx.step_graph_stats()
    return {"graph_iter_per_s": round(rates[True][1] / rates[True][0], 2),
            "eager_iter_per_s": round(rates[False][1] / rates[False][0], 2),
            "windows": f"{windows} x {steps} steps per mode, alternating, steps {tr.step_count - 2 * windows * (steps + 1) + 1}"
                       f"-{tr.step_count}",
            "captures": c1[0] - c0[0], "replays": c1[1] - c0[1], "reruns": c1[2] - c0[2]}


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


def bench_batched(steps=8, warmup=2, seed=10, max_factor=6):
    """Batched views: 1,000,000 SH3 Gaussians (street recipe), 16 views per launch drawn
    from a seeded pool of 32 (pose on the street path, resolution uniform over 640x360,
    1264x832, 1920x1080, 2560x1440; fx = 900 W / 1264).

    Both forms train their own copy of the scene from the same initialisation on the
    same view sequence: the batched form one msplat_hip_train_step_batch per 16 views
    (replayed from its captured graph; gradients summed over the views, one Adam step),
    the plain form one msplat_hip_train_step per view. Views/s compares equal numbers of
    views. Quality is the mean PSNR over the 32 pooled views of each trained model
    (msplat_hip_image_metrics against the targets). The batched form then keeps training,
    outside the views/s window, until its PSNR is within 0.05 dB of the plain form's (at
    most max_factor times the views): the ratio at equal PSNR compares the training time
    each form needs for that quality. Rendering compares msplat_hip_render_batch with 16
    msplat_hip_render on the same views.

    The batched form's Adam follows the batch-size scaling of Zhao et al., "On Scaling Up
    3D Gaussian Splatting Training" (ICLR 2025): with V views per step the learning rates
    scale by sqrt(V) and the moment decays become beta1^V and beta2^V, so that one step
    on V views' summed gradient moves the parameters about as far as V one-view steps."""
    import ctypes as C

    from msplat._abi import Context, DeviceArray, check, lib
    from msplat.backend import (BatchViews, DeviceGaussians, adam_coeffs, render, render_batch, render_numpy,
                                train_step, train_step_batch)
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
    total_steps = warmup + steps * max_factor
    picks = [rng.permutation(BATCH_POOL)[:BATCH_VIEWS] for _ in range(total_steps)]
    mpx = [sum(pool[i]["W"] * pool[i]["H"] for i in p) / 1e6 for p in picks]
    metrics = (C.c_double * 3)()

    def pool_psnr(dg):
        out = []
        for cam, tgt in zip(pool, targets):
            img = render(ctx, dg, cam, BG)
            check(lib().msplat_hip_image_metrics(ctx.ptr, C.c_void_p(img), C.c_void_p(tgt.ptr), cam["H"], cam["W"],
                                                 metrics))
            out.append(metrics[0])
        return float(np.mean(out))

    def fresh():
        return DeviceGaussians.from_points(ctx, pts, rgb, degree=3)

    # plain training, one view per launch
    dp = fresh()
    t = 0
    for p in picks[:warmup]:
        for i in p:
            t += 1
            s_, b_ = adam_coeffs(lrs, t)
            train_step(ctx, dp, pool[i], BG, targets[i], s_, b_, read_loss=False)
    ctx.sync()
    tp = time.perf_counter()
    for p in picks[warmup:warmup + steps]:
        for i in p:
            t += 1
            s_, b_ = adam_coeffs(lrs, t)
            train_step(ctx, dp, pool[i], BG, targets[i], s_, b_, read_loss=False)
    ctx.sync()
    plain_s = time.perf_counter() - tp
    psnr_plain = pool_psnr(dp)
    V = BATCH_VIEWS
    lrs_b, b1_b, b2_b = [lr * float(np.sqrt(V)) for lr in lrs], 0.9 ** V, 0.999 ** V
    # batched training, graph replays; every view set is rendered once first so that the
    # captured step's capacities are grown before the timed steps
    db = fresh()
    batches = [BatchViews([pool[i] for i in p], 3, bg, [targets[i] for i in p]) for p in picks]
    for b in batches:
        render_batch(ctx, db, b)
    ctx.set_batch_graph(True)
    for t_, b in enumerate(batches[:warmup], start=1):
        s_, b_ = adam_coeffs(lrs_b, t_, b1_b, b2_b)
        train_step_batch(ctx, db, b, s_, b_, beta1=b1_b, beta2=b2_b)
    ctx.sync()
    cap0, rep0 = ctx.batch_graph_stats()
    per_step = []
    for t_, b in enumerate(batches[warmup:warmup + steps], start=warmup + 1):
        s_, b_ = adam_coeffs(lrs_b, t_, b1_b, b2_b)
        tb = time.perf_counter()
        train_step_batch(ctx, db, b, s_, b_, beta1=b1_b, beta2=b2_b)  # syncs (reads the overflow flag)
        per_step.append(time.perf_counter() - tb)
    cap1, rep1 = ctx.batch_graph_stats()
    batched_s = float(np.sum(per_step))
    psnr_batched = pool_psnr(db)
    # keep training the batched form until it matches the plain form's quality
    target = psnr_plain - 0.05
    reached, psnr_now, k = (batched_s if psnr_batched >= target else None), psnr_batched, warmup + steps
    while reached is None and k < total_steps:
        s_, b_ = adam_coeffs(lrs_b, k + 1, b1_b, b2_b)
        tb = time.perf_counter()
        train_step_batch(ctx, db, batches[k], s_, b_, beta1=b1_b, beta2=b2_b)
        per_step.append(time.perf_counter() - tb)
        k += 1
        psnr_now = pool_psnr(db)
        if psnr_now >= target:
            reached = float(np.sum(per_step))
    # rendering: batched vs plain on the same views
    bvs = [BatchViews([pool[i] for i in p], 3, bg) for p in picks[:warmup + steps]]
    for b in bvs[:warmup]:
        render_batch(ctx, dp, b)
    ctx.sync()
    tr0 = time.perf_counter()
    for b in bvs[warmup:]:
        render_batch(ctx, dp, b)
    ctx.sync()
    rb_s = time.perf_counter() - tr0
    tr1 = time.perf_counter()
    for p in picks[warmup:warmup + steps]:
        for i in p:
            render(ctx, dp, pool[i], BG)
    ctx.sync()
    rp_s = time.perf_counter() - tr1
    views = BATCH_VIEWS * steps
    out = {
        "workload": "1,000,000 SH3 Gaussians, 16 views per launch from a pool of 32 (resolutions uniform over "
                    "640x360/1264x832/1920x1080/2560x1440)",
        "mean_mpx_per_launch": round(float(np.mean(mpx[warmup:warmup + steps])), 2),
        "train_views_per_s": {"batched": round(views / batched_s, 2), "plain": round(views / plain_s, 2),
                              "ratio": round(plain_s / batched_s, 3)},
        "psnr_after_equal_views_db": {"batched": round(psnr_batched, 4), "plain": round(psnr_plain, 4),
                                      "delta": round(psnr_batched - psnr_plain, 4)},
        "equal_psnr": {"batched_steps_needed": (k - warmup) if reached is not None else None,
                       "batched_psnr_db": round(psnr_now, 4),
                       "ratio": round(plain_s / reached, 3) if reached is not None else None,
                       "note": "training time of the plain form's views / time the batched form needs to come "
                               "within 0.05 dB of its PSNR" if reached is not None else
                               f"the batched form stayed below within {max_factor}x the views"},
        "render_views_per_s": {"batched": round(views / rb_s, 2), "plain": round(views / rp_s, 2),
                               "ratio": round(rp_s / rb_s, 3)},
        "graph": {"captures_in_timed_steps": cap1 - cap0, "replays_in_timed_steps": rep1 - rep0,
                  "captures_total": cap1},
        "adam_scaling": {"lr_factor": round(float(np.sqrt(V)), 3), "beta1": round(b1_b, 6), "beta2": round(b2_b, 6)},
        "launches_timed": steps, "setup_s": round(tp - t0, 1),
    }
    ctx.close()
    return out


SERVE_SCENES = (500_000, 1_000_000, 2_000_000, 3_000_000)
SERVE_RATES = (200, 500, 1000, 2000)


def serve_runtime_scene(srv, ctx, path, centers, seed, trained_steps=None, rate=1000.0, duration=0.6,
                        evict_frac=0.6):
    """A trained scene loaded from its PLY file is admitted into the running server beside
    the resident scenes (Server.add_scene), served in the mix of a Poisson burst, and
    evicted part-way through it (Server.remove_scene, from another thread): its later
    requests are refused and its memory returns to the slots."""
    import threading
    from msplat.scene_io import load_ply, ply_info
    from msplat.serving import latency_stats, poisson_requests
    t0 = time.perf_counter()
    info = ply_info(path)
    scene, _ = load_ply(ctx, path, capacity=info["num_points"])
    ctx.sync()
    load_s = time.perf_counter() - t0
    center = tuple(float(v) for v in scene.params["means"].numpy()[:scene.n].mean(axis=0))
    before, nbytes = srv.resident, scene.nbytes
    idx = srv.add_scene(scene)
    out = {"file": "PLY written from the benchmark's trained street scene", "gaussians": scene.n,
           "trained_steps": trained_steps, "bytes": nbytes, "load_s": round(load_s, 2), "admitted": idx is not None}
    if idx is None:
        scene.free()
        return out
    reqs = poisson_requests(rate, duration, list(centers) + [center], seed=seed)
    for r in reqs:
        if r.scene == len(centers):
            r.scene = idx
    timer = threading.Timer(evict_frac * duration, srv.remove_scene, args=(idx,))
    timer.start()
    done = srv.run(reqs)
    timer.join()
    mine = [r for r in done if r.scene == idx]
    refused = [r for r in srv.rejected if r.scene == idx]
    stats = latency_stats(mine)
    out.update({"index": idx, "resident_gb": [round(before / 1e9, 2), round((before + nbytes) / 1e9, 2),
                                              round(srv.resident / 1e9, 2)],
                "rate": rate, "duration_s": duration, "evicted_at_s": round(evict_frac * duration, 2),
                "requests": sum(r.scene == idx for r in reqs), "served": len(mine),
                "refused_after_eviction": len(refused), "p50_ms": stats.get("p50_ms"), "p99_ms": stats.get("p99_ms"),
                "others": {k: latency_stats([r for r in done if r.scene != idx]).get(k)
                           for k in ("completed", "p50_ms", "p99_ms")}})
    return out


SHED_WAIT_MS = 4.0  # queueing bound of the load-shedding run


def bench_serving(rates=SERVE_RATES, duration=1.0, slots=6, max_views=16, seed=20, trained_ply=None,
                  trained_steps=None):
    """Serving: four resident SH3 scenes (0.5M, 1M, 2M, 3M Gaussians, 64-cluster clouds in
    an 8 m box, kNN scales, opacity logits N(1.5, 1), higher SH bands N(0, 0.05)).
    Poisson arrivals at each rate for `duration` s; pose on a sphere of radius 3-10 m
    around the scene's centre, resolution 640x360 40% / 1280x720 30% / 1920x1080 20% /
    3840x2160 10%. The continuous-batching server (6 slots, up to 16 views per batch, an
    HBM budget of the resident scenes plus 90% of the device's free memory) against the
    plain form (one msplat_hip_render per request, first come first served), each
    followed by the RGBA8 pack. A last run at the top rate gives the server a budget that
    fits two 1920x1080 working sets per slot but no 3840x2160 one, to show admission
    refusing what cannot fit; another sheds the requests still queued SHED_WAIT_MS after
    their arrival, to show the latency of what is served under overload."""
    from msplat._abi import Context
    from msplat.backend import DeviceGaussians
    from msplat.scenes import gmm_cloud
    from msplat.serving import Request, Server, latency_stats, poisson_requests, serve_plain
    ctx = Context(0)
    t0 = time.perf_counter()
    rng = np.random.default_rng(seed)
    centers = [(60.0 * k, 0.0, 0.0) for k in range(len(SERVE_SCENES))]
    scenes = []
    for k, (n, c) in enumerate(zip(SERVE_SCENES, centers)):
        pts, rgb = gmm_cloud(seed + k, n, c, box=8.0)
        dg = DeviceGaussians.from_points(ctx, pts, rgb, degree=3)
        dg.params["opac"].upload(rng.normal(1.5, 1.0, (n, 1)).astype(np.float32))
        dg.params["rest"].upload(rng.normal(0, 0.05, (n, 45)).astype(np.float32))
        scenes.append(dg)
    ctx.sync()
    setup = time.perf_counter() - t0
    srv = Server(scenes, slots=slots, max_views=max_views, background=BG)
    warm = poisson_requests(2000.0, 0.1, centers, seed=seed + 99)
    srv.run(warm)
    serve_plain(scenes, poisson_requests(500.0, 0.1, centers, seed=seed + 98), background=BG)
    out = {"workload": "4 resident SH3 scenes (0.5M/1M/2M/3M Gaussians); Poisson arrivals; 640x360 40% / "
                       "1280x720 30% / 1920x1080 20% / 3840x2160 10%; pose radius 3-10 m",
           "slots": slots, "max_views": max_views, "duration_s": duration, "setup_s": round(setup, 1),
           "hbm_budget_gb": round(srv.budget / 1e9, 1), "resident_gb": round(srv.resident / 1e9, 2), "rates": {}}
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
    if trained_ply:
        out["runtime_scene"] = serve_runtime_scene(srv, ctx, trained_ply, centers, seed + 7, trained_steps)
        log("runtime scene: " + json.dumps(out["runtime_scene"]))
    srv.close()
    # admission under a tight budget at the top rate
    probe = Server(scenes, slots=1, background=BG)
    ipp = max(probe.ipp)
    fhd = Request(0, len(scenes) - 1, {"W": 1920, "H": 1080}, 0.0)
    tight = probe.resident + slots * 2.0 * probe.working_set([fhd])
    probe.close()
    srv = Server(scenes, slots=slots, max_views=max_views, background=BG, hbm_budget=tight, isects_per_pixel=ipp,
                 adapt=False)
    reqs = poisson_requests(float(rates[-1]), duration, centers, seed=seed + 7)
    b = latency_stats(srv.run(reqs))
    rej = {}
    for r in srv.rejected:
        k = f"{r.cam['W']}x{r.cam['H']}"
        rej[k] = rej.get(k, 0) + 1
    out["admission"] = {"rate": rates[-1], "hbm_budget_gb": round(tight / 1e9, 2), "requests": len(reqs),
                        "completed": b.get("completed", 0), "rejected_by_resolution": rej,
                        "p99_ms": b.get("p99_ms"), "views_per_s": b.get("views_per_s")}
    log(f"admission at a {tight / 1e9:.2f} GB budget: rejected {rej}")
    srv.close()
    # load shedding at the top rate: requests queued longer than max_wait_ms are dropped
    srv = Server(scenes, slots=slots, max_views=max_views, background=BG, max_wait_ms=SHED_WAIT_MS)
    reqs = poisson_requests(float(rates[-1]), duration, centers, seed=seed + rates[-1])
    b = latency_stats(srv.run(reqs))
    shed = {}
    for r in srv.shed:
        k = f"{r.cam['W']}x{r.cam['H']}"
        shed[k] = shed.get(k, 0) + 1
    out["shedding"] = {"rate": rates[-1], "max_wait_ms": SHED_WAIT_MS, "requests": len(reqs),
                       "completed": b.get("completed", 0), "shed": len(srv.shed), "shed_by_resolution": shed,
                       "views_per_s": b.get("views_per_s"), "p50_ms": b.get("p50_ms"), "p99_ms": b.get("p99_ms"),
                       "p99_1080p_ms": b.get("by_resolution", {}).get("1920x1080", {}).get("p99_ms")}
    log("shedding: " + json.dumps(out["shedding"]))
    srv.close()
    ctx.close()
    return out


def tune_train_step(ctx, cams, targets, seed=5):
    """Picks the rasterizer variants (msplat.autotune) on the headline train step, run on
    a scratch copy of the scene (the same street recipe) so the trainer's schedule is not
    advanced; the choice stays set on the context."""
    import ctypes as C
    from msplat import autotune
    from msplat._abi import check, lib
    from msplat.backend import KEYS, DeviceGaussians, adam_coeffs, train_step
    from msplat.scenes import street_points
    pts, rgb = street_points(seed, N_GAUSSIANS)
    scratch = DeviceGaussians.from_points(ctx, pts, rgb, degree=3)
    start = {k: ctx.download(scratch.params[k].ptr, (scratch.params[k].nbytes,), np.uint8) for k in KEYS}
    lrs = [1.6e-4, 5e-3, 1e-3, 2.5e-3, 1.25e-4, 5e-2]
    k = [0]

    def step():
        k[0] += 1
        s_, b_ = adam_coeffs(lrs, k[0])
        i = k[0] % len(cams)
        train_step(ctx, scratch, dict(cams[i], dtu=3), BG, targets[i], s_, b_, read_loss=False)

    def reset():  # every trial trains the same first steps of the same scene
        for key in KEYS:
            ctx.upload(scratch.params[key].ptr, start[key])
            for m in (scratch.m1[key], scratch.m2[key]):
                check(lib().msplat_hip_memset(ctx.ptr, C.c_void_p(m.ptr), 0, m.nbytes))
        scratch.reset_stats()
        k[0] = 0

    # two rounds: the best pixels per lane depends on the batch depth picked after it
    res = autotune.autotune(ctx, step, warmup=3, repeats=24, rounds=2, reset=reset)
    log("tuning trials (ms per step): " + "; ".join(
        "%s %.4f" % ("/".join(str(c[n]) for n in autotune.SPACE), ms) for c, ms in res["trials"]))
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


def kernel_key(name: str) -> str:
    """'void msplat::raster_backward_kernel<2, 64, false>(msplat::RasterArgs)' ->
    'raster_backward_kernel<2, 64, false>'."""
    n = name.strip().strip('"')
    if n.startswith("void "):
        n = n[5:]
    n = n.replace("msplat::", "")
    depth, cut = 0, len(n)
    for i, ch in enumerate(n):  # drop the argument list (the first '(' outside template brackets)
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            cut = i
            break
    return n[:cut]


def bwd_kernel_name(ppt: int, bb: int) -> str:
    """The one-view raster backward's demangled name (template: pixels per thread, LDS
    batch depth, prior-guided aux sums off)."""
    return "raster_backward_kernel<%d, %d, false>" % (ppt, bb)


def profiled_counters(variant: dict, warmup: int, steps: int):
    """Per-launch counters of the raster backward variant `variant` on the timed steps,
    from the committed counter summary (rocprofv3 --pmc runs of the same workload and
    window, HBM counters scaled by a copy of known size); None when absent or when the
    summary's step window is not this run's (a launch's work changes over the schedule:
    the splats shrink as training goes on)."""
    if not os.path.exists(COUNTERS):
        return None
    prof = json.load(open(COUNTERS))
    if prof.get("window") != {"warmup": warmup, "steps": steps}:
        return None
    name = bwd_kernel_name(variant.get("bwd_ppt", 4), variant.get("bwd_batch", 64))
    entry = prof.get("kernels", {}).get(name)
    return dict(entry, kernel=name) if entry else None


class _Roctx:
    """roctx ranges (no-ops unless a profiler with marker tracing is attached)."""

    def __init__(self):
        import ctypes as C
        self.lib = None
        for lib in ("librocprofiler-sdk-roctx.so.1", "libroctx64.so.4"):
            try:
                self.lib = C.CDLL(lib)
                self.lib.roctxRangePushA.argtypes = [C.c_char_p]
                break
            except OSError:
                continue

    def push(self, name: str):
        if self.lib:
            self.lib.roctxRangePushA(name.encode())

    def pop(self):
        if self.lib:
            self.lib.roctxRangePop()


def _csv_rows(pattern_dir, suffix):
    import csv
    import glob
    out = []
    for f in sorted(glob.glob(os.path.join(pattern_dir, "**", "*" + suffix), recursive=True)):
        with open(f) as fh:
            out.extend(csv.DictReader(fh))
    return out


def summarize_profile(dirs, warmup, steps, calibration=None, anchor="raster_backward_kernel"):
    """Summarises rocprofv3 output directories of this benchmark: kernel traces (the
    timed steps are the 'timed_steps' roctx range) and --pmc counter runs (the timed
    steps are the dispatches after the warmup-th and up to the (warmup+steps)-th
    dispatch of the anchor kernel, which runs once per step). FETCH_SIZE / WRITE_SIZE
    are scaled by the copy calibration (bench.py --copy-calibration: 256 MiB read and
    written per dispatch). Prints JSON."""
    res = {"window": {"warmup": warmup, "steps": steps}, "kernels": {}}
    scale = {"FETCH_SIZE": 1.0, "WRITE_SIZE": 1.0}
    for cal in calibration or ():
        for row in _csv_rows(cal, "counter_collection.csv"):
            if "copyBuffer" in row["Kernel_Name"] and row["Counter_Name"] in scale:
                scale[row["Counter_Name"]] = 262144.0 / float(row["Counter_Value"])
        res["fetch_scale"], res["write_scale"] = round(scale["FETCH_SIZE"], 4), round(scale["WRITE_SIZE"], 4)
    for d in dirs:
        trace = _csv_rows(d, "kernel_trace.csv")
        if trace:
            marks = [r for r in _csv_rows(d, "marker_api_trace.csv") if r.get("Function") == "timed_steps"]
            if not marks:
                raise SystemExit(f"{d}: no 'timed_steps' range in the marker trace")
            t0, t1 = int(marks[0]["Start_Timestamp"]), int(marks[0]["End_Timestamp"])
            per = {}
            for r in trace:
                a, b = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
                if t0 <= a and b <= t1:
                    per.setdefault(kernel_key(r["Kernel_Name"]), []).append(b - a)
            for k, v in per.items():
                e = res["kernels"].setdefault(k, {})
                e.update(dispatches_per_step=round(len(v) / steps, 3), avg_us=round(sum(v) / len(v) / 1e3, 3),
                         us_per_step=round(sum(v) / steps / 1e3, 3))
            res["timed_window_ms"] = round((t1 - t0) / 1e6, 3)
            res["kernel_us_per_step"] = round(sum(e.get("us_per_step", 0) for e in res["kernels"].values()), 1)
            continue
        rows = _csv_rows(d, "counter_collection.csv")
        if not rows:
            continue
        by_id = {}
        for r in rows:
            e = by_id.setdefault(int(r["Dispatch_Id"]), {"name": kernel_key(r["Kernel_Name"])})
            e[r["Counter_Name"]] = float(r["Counter_Value"])
        ids = sorted(by_id)
        anchors = [i for i in ids if by_id[i]["name"].startswith(anchor)]
        if len(anchors) < warmup + steps:
            raise SystemExit(f"{d}: {len(anchors)} dispatches of {anchor}, fewer than warm-up + timed steps")
        lo, hi = anchors[warmup - 1] if warmup else -1, anchors[warmup + steps - 1]
        acc = {}
        for i in ids:
            if lo < i <= hi:
                acc.setdefault(by_id[i]["name"], []).append(by_id[i])
        for k, lst in acc.items():
            e = res["kernels"].setdefault(k, {})
            for c in lst[0]:
                if c == "name":
                    continue
                mean = sum(x.get(c, 0.0) for x in lst) / len(lst)
                if c in scale:
                    e[c.lower() + "_bytes"] = round(mean * 1024 * scale[c])
                else:
                    e[c] = round(mean, 1)
            e["dispatches_per_step_pmc"] = round(len(lst) / steps, 3)
    for k, e in res["kernels"].items():
        if "fetch_size_bytes" in e and "write_size_bytes" in e:
            e["hbm_bytes"] = e["fetch_size_bytes"] + e["write_size_bytes"]
        if "SQ_INSTS_VALU" in e and "avg_us" in e:
            e["valu_issue_util"] = round(e["SQ_INSTS_VALU"] * 4 / (SIMDS * CLOCK_HZ * e["avg_us"] * 1e-6), 4)
        if "SQ_ACTIVE_INST_VALU" in e and e.get("GRBM_GUI_ACTIVE"):
            e["VALUBusy_pct"] = round(100 * e["SQ_ACTIVE_INST_VALU"] / 256 / e["GRBM_GUI_ACTIVE"], 2)
        if "SQ_LDS_IDX_ACTIVE" in e and e.get("GRBM_GUI_ACTIVE"):
            e["LdsUtil_pct"] = round(100 * e["SQ_LDS_IDX_ACTIVE"] / 256 / e["GRBM_GUI_ACTIVE"], 2)
        if "SQ_LDS_BANK_CONFLICT" in e and e.get("GRBM_GUI_ACTIVE"):
            e["LDSBankConflict_pct"] = round(100 * e["SQ_LDS_BANK_CONFLICT"] / 256 / e["GRBM_GUI_ACTIVE"], 2)
    print(json.dumps(res, indent=1, sort_keys=True))


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def cpu_baseline_sample(dg, cams, targets, step_index, seconds=20.0, plain=None):
    """Times the reference's CPU train step (core/cpu/msplat_cpu.cpp compiled into
    oracle/_ref, optimised build for the widest x86-64 level the host runs) — or the
    oracle when the reference was not built — on this host's cores (MSPLAT_THREADS = the
    CPUs this process may run on), on the workload's own scene and first view at the SH
    degree the timed steps use."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    threads = len(os.sched_getaffinity(0))
    os.environ["MSPLAT_THREADS"] = str(threads)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    from oracle import Oracle, Reference, host_isa, reference_available
    isa = host_isa()
    kind = "reference" if reference_available(perf=True, isa=isa) else "port"
    impl = Reference(perf=True, isa=isa) if kind == "reference" else Oracle()
    full = dg.to_numpy()
    keys = ("means", "scales", "quats", "dc", "rest", "opac")
    st = {k: full[k].copy() for k in keys}
    for k in keys:
        st["m1_" + k] = np.zeros_like(full[k])
        st["m2_" + k] = np.zeros_like(full[k])
    n = full["means"].shape[0]
    st.update(degree=dg.degree, vis=np.zeros(n, np.float32), gn=np.zeros(n, np.float32), ms=np.zeros(n, np.float32))
    dtu = min(step_index // 1000, 3)
    cam = dict(cams[0], dtu=dtu)
    gt = targets[0].numpy()
    lrs = np.array([1.6e-4, 5e-3, 1e-3, 2.5e-3, 1.25e-4, 5e-2], np.float32)
    steps, bc2 = lrs / np.float32(0.1), np.full(6, np.sqrt(np.float32(0.001)), np.float32)
    times = []
    t_start = time.perf_counter()
    while True:
        t0 = time.perf_counter()
        impl.train_step(st, cam, np.array(BG, np.float32), gt, steps, bc2)
        times.append(time.perf_counter() - t0)
        if time.perf_counter() - t_start > seconds or len(times) >= 5:
            break
    per = float(np.median(times))
    build = f"-O3 -march=x86-64-{isa}" if kind == "reference" else "-O2 (oracle)"
    plain_cpu = cpu_plain_run(plain, impl, threads) if plain else None
    return {"value": round(1.0 / per, 4), "unit": "iter/s", "cores": threads, "kind": kind, "plain": plain_cpu,
            "host_cpus": os.cpu_count(), "cpu_model": cpu_model(), "build": build, "sh_degree_in_use": dtu,
            "sample": f"{len(times)} train step(s) of the timed {n:,}-Gaussian scene on view 0 at {W_IMG}x{H_IMG} "
                      f"(SH degree in use {dtu}, as in the timed steps), median {per:.3f} s/step, "
                      f"MSPLAT_THREADS={threads}, {cpu_model()}"}


PLAIN_STEPS = 300


def plain_workload(ctx):
    """The CPU-runnable plain workload: 20,000 Gaussians (means U[-1,1]^3, log-scales
    N(-4.5, 0.5), uniform rotations, logits N(0,1), SH degree 1), 8 pinhole views at
    128x128 (fx = fy = 120) on a radius-4 ring, targets rendered from a second seeded
    20,000-Gaussian set; 300 Adam steps of 0.8 L1 + 0.2 D-SSIM, no densification."""
    from msplat._abi import DeviceArray
    from msplat.backend import DeviceGaussians, render_numpy
    from msplat.scenes import random_scene, ring_cameras
    sc = random_scene(1000, 20000, degree=1)
    cams = ring_cameras(8, 4.0, 128, 128, 120.0, 120.0, dtu=1)
    gt = DeviceGaussians(ctx, random_scene(1001, 20000, degree=1))
    targets = [DeviceArray.from_numpy(ctx, render_numpy(ctx, gt, c, BG)) for c in cams]
    return sc, cams, targets


def bench_plain(repeats=3):
    """The plain workload's whole 300-step run on the GPU (best of `repeats` after one
    warm-up run), with the per-tile rasterizers, with depth chunks chosen per frame (64
    tiles per view: every frame is chunked), and with those steps replayed from a captured
    graph."""
    from msplat._abi import TUNE_CHUNKS, Context
    from msplat.backend import DeviceGaussians
    from msplat.trainer import TrainConfig, Trainer
    ctx = Context(0)
    out = {"workload": "20,000 Gaussians SH1, 8 views at 128x128, 300 Adam steps, no densification"}
    try:
        sc, cams, targets = plain_workload(ctx)
        for name, chunks, graph in (("per_tile", 1, False), ("depth_chunked", 0, False),
                                    ("depth_chunked_step_graph", 0, True)):
            ctx.set_tuning(TUNE_CHUNKS, chunks)
            best, loss = None, None
            for r in range(repeats + 1):
                dg = DeviceGaussians(ctx, sc)
                cfg = TrainConfig(iterations=PLAIN_STEPS, sh_degree=1, sh_degree_interval=1, densify=False, seed=42,
                                  step_graph=graph)
                tr = Trainer(ctx, dg, cams, targets, cfg)
                ctx.sync()
                t0 = time.perf_counter()
                for _ in range(PLAIN_STEPS - 1):
                    tr.step()
                loss = tr.step(read_loss=True)
                el = time.perf_counter() - t0
                if r:
                    best = el if best is None else min(best, el)
            out[name] = {"iter_per_s": round(PLAIN_STEPS / best, 1), "ms_per_step": round(1e3 * best / PLAIN_STEPS, 4),
                         "run_s": round(best, 4), "final_loss": round(float(loss), 6),
                         "depth_chunks": ctx.frame().depth_chunks}
        out["targets"] = [t.numpy() for t in targets]
    finally:
        ctx.close()
    return out


def cpu_plain_run(plain, impl, threads):
    """The plain workload's whole 300-step run through the reference's CPU train step
    (the same schedule: Adam bias corrections per step, means LR decayed over the run),
    on the targets the GPU leg rendered."""
    from msplat.backend import adam_coeffs
    from msplat.scenes import random_scene, ring_cameras
    from msplat.trainer import LR_INIT, means_lr
    sc = random_scene(1000, 20000, degree=1)
    cams = ring_cameras(8, 4.0, 128, 128, 120.0, 120.0, dtu=1)
    keys = ("means", "scales", "quats", "dc", "rest", "opac")
    st = {k: np.ascontiguousarray(sc[k], np.float32).copy() for k in keys}
    for k in keys:
        st["m1_" + k] = np.zeros_like(st[k])
        st["m2_" + k] = np.zeros_like(st[k])
    n = st["means"].shape[0]
    st.update(degree=1, vis=np.zeros(n, np.float32), gn=np.zeros(n, np.float32), ms=np.zeros(n, np.float32))
    rng = np.random.default_rng(42)
    order = []
    while len(order) < PLAIN_STEPS:
        order.extend(int(i) for i in rng.permutation(len(cams)))
    lrs, times, loss = list(LR_INIT), [], None
    for t in range(1, PLAIN_STEPS + 1):
        steps, bc2 = adam_coeffs(lrs, t)
        i = order[t - 1]
        t0 = time.perf_counter()
        loss = impl.train_step(st, cams[i], np.array(BG, np.float32), plain["targets"][i], steps, bc2)[0]
        times.append(time.perf_counter() - t0)
        lrs[0] = means_lr(t, PLAIN_STEPS)
    total = float(np.sum(times))
    return {"value": round(PLAIN_STEPS / total, 2), "unit": "iter/s", "cores": threads, "run_s": round(total, 3),
            "median_step_ms": round(1e3 * float(np.median(times)), 3), "final_loss": round(float(loss), 6),
            "sample": f"the whole {PLAIN_STEPS}-step run"}


def time_sh3_steps(ctx, dg, cams, targets, steps=100, warmup=10):
    """Train steps with all four SH bands in use (degree 3), on the current scene and the
    same views: the rate the 3000-step schedule never reaches (its degree ramp reaches 3
    at step 3000)."""
    from msplat.trainer import TrainConfig, Trainer
    cfg = TrainConfig(iterations=steps + warmup, sh_degree=3, sh_degree_interval=1, densify=False)
    tr = Trainer(ctx, dg, cams, targets, cfg)
    for _ in range(warmup):
        tr.step()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        tr.step()
    ctx.sync()
    return steps / (time.perf_counter() - t0)


STREET_PRIORS = dict(depth_weight=0.2, depth_weight_final=0.05, sky_alpha_weight=0.05, fill_weight=0.01,
                     use_masks=True, learn_sky=True, exposure=True, max_scale_ratio=10.0)


def street_camera_priors(cams, seed=30):
    """Per view, a depth prior and an aux image (confidence, sky mask, keep mask) at half
    the view's resolution, shaped like a forward-looking street capture: sky above the
    horizon row the camera's pitch puts it on (no depth there), a ground plane 1.6 m
    below the camera below it (depth = 1.6 f / rows below the horizon, at most 80 m), 2%
    of the depth samples invalid, and the bottom 6% of rows (the vehicle's hood) not kept."""
    rng = np.random.default_rng(seed)
    out = []
    for c in cams:
        pw, ph = c["W"] // 2, c["H"] // 2
        fy, horizon = 0.5 * c["fy"], 0.5 * (c["cy"] - 0.1 * c["fy"])
        yy = np.mgrid[0:ph, 0:pw][0].astype(np.float32)
        below = yy - np.float32(horizon)
        depth = np.where(below > 0.5, np.minimum(1.6 * fy / np.maximum(below, 0.5), 80.0), 0.0).astype(np.float32)
        depth[rng.random((ph, pw)) < 0.02] = 0.0
        aux = np.zeros((ph, pw, 4), np.uint8)
        aux[..., 0] = rng.integers(120, 256, (ph, pw))
        aux[..., 1] = np.where(below < -4.0, 255, 0)
        aux[..., 2] = np.where(yy < 0.94 * ph, 255, 0)
        out.append(dict(depth=depth, aux=aux, has_depth=True, has_sky=True, has_mask=True))
    return out


def time_prior_steps(ctx, dg, cams, targets, steps=100, warmup=10):
    """Prior-guided train steps with the street preset (python/msplat/__init__.py:61-72:
    depth prior 0.2 -> 0.05, sky alpha 0.05, fill 0.01, keep masks, learned 512x128 sky,
    per-view exposure, scale cap 10; SH degree 1) on the current scene and views, with
    street_camera_priors; and, beside them, plain steps at SH degree 1 on the same scene
    for the prior path's cost. Returns the rates and what the CPU leg needs."""
    from msplat.trainer import PriorOptions, TrainConfig, Trainer
    cps = street_camera_priors(cams)
    rates = {}
    for name, priors in (("plain_sh1", None), ("prior", PriorOptions(**STREET_PRIORS))):
        cfg = TrainConfig(iterations=steps + warmup, sh_degree=1, sh_degree_interval=1, densify=False)
        tr = Trainer(ctx, dg, cams, targets, cfg, priors=priors, camera_priors=cps if priors else None)
        for _ in range(warmup):
            tr.step()
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            tr.step()
        ctx.sync()
        rates[name] = steps / (time.perf_counter() - t0)
    sky = tr.prior_state.numpy("sky_tex")
    return {"iter_per_s": round(rates["prior"], 2), "plain_sh1_iter_per_s": round(rates["plain_sh1"], 2),
            "prior_cost_pct": round(100.0 * (rates["plain_sh1"] / rates["prior"] - 1.0), 1),
            "workload": f"street preset on the timed scene ({dg.n:,} Gaussians at the end), one {W_IMG}x{H_IMG} "
                        f"view per step, priors at {W_IMG // 2}x{H_IMG // 2} (synthetic ground-plane depth, sky "
                        f"mask above the horizon, hood not kept), {steps} steps after {warmup}, no densification",
            "_cpu": dict(sky=sky, sky_frame=tr.sky_frame, prior=cps[0], n_cams=len(cams))}


def cpu_prior_sample(dg, cams, targets, leg, seconds=10.0):
    """The reference's CPU prior-guided train step (ref_train_step_prior, the compiled
    core/cpu/msplat_cpu.cpp) on the same scene, view 0 and its priors, the sky texture the
    GPU leg ended with and the preset's weights at the first step; bounded to `seconds`
    or 5 steps. None when the reference was not built."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Reference, host_isa, reference_available
    isa = host_isa()
    if not reference_available(perf=True, isa=isa):
        return None
    threads = len(os.sched_getaffinity(0))
    os.environ["MSPLAT_THREADS"] = str(threads)
    impl = Reference(perf=True, isa=isa)
    full = dg.to_numpy()
    keys = ("means", "scales", "quats", "dc", "rest", "opac")
    st = {k: full[k].copy() for k in keys}
    for k in keys:
        st["m1_" + k] = np.zeros_like(full[k])
        st["m2_" + k] = np.zeros_like(full[k])
    n = full["means"].shape[0]
    st.update(degree=dg.degree, vis=np.zeros(n, np.float32), gn=np.zeros(n, np.float32), ms=np.zeros(n, np.float32))
    c, o = leg["_cpu"], STREET_PRIORS
    sky = np.ascontiguousarray(c["sky"], np.float32)
    ident = np.zeros((c["n_cams"], 12), np.float32)
    ident[:, [0, 4, 8]] = 1.0
    bc2 = float(np.sqrt(np.float32(1 - 0.999)))
    prior = dict(aux=True, sky=True, sky_tex=sky, sky_grad=np.zeros_like(sky), sky_m1=np.zeros_like(sky),
                 sky_m2=np.zeros_like(sky), sky_frame=tuple(float(v) for v in c["sky_frame"]),
                 sky_step=0.01 / 0.1, sky_bc2=bc2, prior_depth=c["prior"]["depth"], prior_aux=c["prior"]["aux"],
                 has_depth=True, has_sky_mask=True, depth_weight=o["depth_weight"], sky_weight=o["sky_alpha_weight"],
                 fill_weight=o["fill_weight"], huber_delta=0.05, min_alpha=0.25, mask_photometric=True, exposure=True,
                 cam_index=0, expo_params=ident, expo_m1=np.zeros_like(ident), expo_m2=np.zeros_like(ident),
                 expo_grad=np.zeros(12, np.float32), expo_step=5e-3 / 0.1, expo_bc2=bc2, expo_reg=1e-2,
                 log_max_scale_ratio=float(np.log(np.float32(o["max_scale_ratio"]))), loss_terms=np.zeros(4, np.float32))
    lrs = np.array([1.6e-4, 5e-3, 1e-3, 2.5e-3, 1.25e-4, 5e-2], np.float32)
    steps, bc2s = lrs / np.float32(0.1), np.full(6, np.sqrt(np.float32(0.001)), np.float32)
    times, t_start = [], time.perf_counter()
    while True:
        t0 = time.perf_counter()
        impl.train_step_prior(st, dict(cams[0], dtu=1), np.array(BG, np.float32), targets[0].numpy(), steps, bc2s,
                              prior)
        times.append(time.perf_counter() - t0)
        if time.perf_counter() - t_start > seconds or len(times) >= 5:
            break
    per = float(np.median(times))
    return {"value": round(1.0 / per, 4), "unit": "iter/s", "cores": threads, "kind": "reference",
            "build": f"-O3 -march=x86-64-{isa}",
            "sample": f"{len(times)} prior-guided train step(s) on view 0 (SH degree 1 in use), "
                      f"median {per:.3f} s/step, MSPLAT_THREADS={threads}"}


def bench_images(repeats=200, cpu_seconds=5.0, cpu=True):
    """The training images' pipeline for one 1264x832 view: area resampling of a decoded
    RGBA8 camera image (2528x1664, and a 4032x3024 one at a non-integer ratio) to the
    working size, one pyramid level (1264x832 -> 632x416 float), and Brown-Conrady
    undistortion into the full frame (k1 -0.08, k2 0.02, p1 1e-3, p2 -5e-4). Per call:
    `repeats` back-to-back launches from the host, wall clock over them (so launch costs
    are included), bytes = source read once + output written once. Beside them the
    oracle's single-thread C restatement (the reference's msplat_resize_area /
    msplat_undistort, core/cpu/msplat_cpu.cpp:1609-1669) on the same inputs, as the
    checker and the CPU baseline (skipped when `cpu` is False)."""
    import ctypes as C
    from msplat._abi import Context, DeviceArray, check, floats, lib
    orc = None
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        from oracle import Oracle
        orc = Oracle()
    ctx = Context(0)
    rng = np.random.default_rng(40)
    intr, dist = (FOCAL, FOCAL, W_IMG / 2.0, H_IMG / 2.0), (-0.08, 0.02, 1e-3, -5e-4, 0.0)
    out = {"view": f"{W_IMG}x{H_IMG}", "ops": {}}
    try:
        cases = []
        for sw, sh in ((2 * W_IMG, 2 * H_IMG), (4032, 3024)):
            img = rng.integers(0, 256, (sh, sw, 4), dtype=np.uint8)
            cases.append((f"resize_u8_{sw}x{sh}", img, W_IMG, H_IMG))
        lvl = rng.uniform(0, 1, (H_IMG, W_IMG, 3)).astype(np.float32)
        cases.append((f"resize_f32_{W_IMG}x{H_IMG}_to_half", lvl, W_IMG // 2, H_IMG // 2))
        cases.append((f"undistort_f32_{W_IMG}x{H_IMG}", lvl, W_IMG, H_IMG))
        for name, host, dw, dh in cases:
            src = DeviceArray.from_numpy(ctx, host, host.dtype)
            dst = DeviceArray(ctx, (dh, dw, 3), np.float32, zero=False)
            sh, sw = host.shape[:2]
            if name.startswith("undistort"):
                fi, fd = floats(intr), floats(dist)

                def call():
                    check(lib().msplat_hip_undistort(ctx.ptr, C.c_void_p(src.ptr), sw, sh, fi, fd, 0, 0,
                                                     C.c_void_p(dst.ptr), dw, dh))
            else:
                u8 = int(host.dtype == np.uint8)

                def call():
                    check(lib().msplat_hip_resize_area(ctx.ptr, C.c_void_p(src.ptr), u8, sw, sh, C.c_void_p(dst.ptr),
                                                       dw, dh))
            for _ in range(10):
                call()
            ctx.sync()
            t0 = time.perf_counter()
            for _ in range(repeats):
                call()
            ctx.sync()
            per = (time.perf_counter() - t0) / repeats
            nbytes = host.nbytes + dw * dh * 12
            out["ops"][name] = {"us_per_call": round(1e6 * per, 2), "bytes": int(nbytes),
                                "gbs": round(nbytes / per / 1e9, 1), "frac": round(nbytes / per / 1e9 / HBM_PEAK_GBS, 4)}
            if orc is not None:
                got = dst.numpy().reshape(dh, dw, 3)
                times, t_start = [], time.perf_counter()
                while True:
                    t1 = time.perf_counter()
                    ref = (orc.undistort(host, intr, dist, 0, 0, dw, dh) if name.startswith("undistort")
                           else orc.resize_area(host, dw, dh))
                    times.append(time.perf_counter() - t1)
                    if time.perf_counter() - t_start > cpu_seconds / len(cases) or len(times) >= 20:
                        break
                out["ops"][name].update(bit_exact_vs_oracle=bool(np.array_equal(got, ref)),
                                        cpu_port_us=round(1e6 * float(np.median(times)), 1))
            src.free()
            dst.free()
    finally:
        ctx.close()
    return out


def bench_scene_files(ctx, dg, step, tmpdir, repeats=3, cpu=True):
    """The trained scene's files (save_gaussians.cpp:10-191, model.cpp:282-496): PLY and
    .splat exports, the PLY import and the .msplat checkpoint, each the best of `repeats`
    (host I/O into `tmpdir` after one device-to-host staging copy, so the page cache and
    the host's write speed bound them); beside them the reference's own writers and reader
    (the compiled core/src/loaders/save_gaussians.cpp) on the same parameters, whose files
    are compared byte for byte. The reference is skipped when it was not built or `cpu` is
    False."""
    from msplat.scene_io import load_checkpoint, load_ply, save_checkpoint, save_ply, save_splat

    def best(fn):
        times = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            fn()
            times.append(time.perf_counter() - t0)
        return min(times)

    p = {k: os.path.join(tmpdir, "scene." + k) for k in ("ply", "splat", "msplat", "ref.ply", "ref.splat")}
    lrs = [1.6e-4, 5e-3, 1e-3, 2.5e-3, 1.25e-4, 5e-2]
    loaded = []

    def load(fn):
        loaded.append(fn())
        loaded.pop()[0].free()

    gpu = {"save_ply": best(lambda: save_ply(ctx, dg, p["ply"], step=step)),
           "save_splat": best(lambda: save_splat(ctx, dg, p["splat"])),
           "save_checkpoint": best(lambda: save_checkpoint(ctx, p["msplat"], dg, step, step, lrs)),
           "load_ply": best(lambda: load(lambda: load_ply(ctx, p["ply"], capacity=dg.n))),
           "load_checkpoint": best(lambda: load(lambda: load_checkpoint(ctx, p["msplat"], capacity=dg.n)))}
    size = {"save_ply": p["ply"], "save_splat": p["splat"], "save_checkpoint": p["msplat"], "load_ply": p["ply"],
            "load_checkpoint": p["msplat"]}
    out = {"gaussians": dg.n, "sh_degree": dg.degree,
           "ops": {k: {"s": round(t, 4), "mb": round(os.path.getsize(size[k]) / 1e6, 1),
                       "mb_per_s": round(os.path.getsize(size[k]) / t / 1e6, 1)} for k, t in gpu.items()}}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        from oracle import Reference, reference_available
    if cpu and reference_available():
        ref = Reference()
        host = dg.to_numpy()
        scene = {k: host[k] for k in ("means", "scales", "quats", "dc", "rest", "opac")}
        scene["degree"] = dg.degree
        rb = dg.rest_bases
        out["reference"] = {
            "save_ply_s": round(best(lambda: ref.save(scene, p["ref.ply"], step=step)), 4),
            "save_splat_s": round(best(lambda: ref.save(scene, p["ref.splat"], splat=True)), 4),
            "load_ply_s": round(best(lambda: ref.load_ply(p["ref.ply"], dg.n, rb)), 4),
            "ply_bytes_equal": open(p["ply"], "rb").read() == open(p["ref.ply"], "rb").read(),
            "splat_bytes_equal": open(p["splat"], "rb").read() == open(p["ref.splat"], "rb").read(),
            "build": "oracle/_ref/libmsplat_ref.so (the reference's -O2 build), one thread"}
    return out


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
    ap.add_argument("--no-sh3", action="store_true", help="skip timing steps with SH degree 3 in use")
    ap.add_argument("--no-prior", action="store_true", help="skip timing prior-guided steps (street preset)")
    ap.add_argument("--no-images", action="store_true", help="skip timing the training images' pipeline")
    ap.add_argument("--no-scene-files", action="store_true", help="skip timing the trained scene's file exports")
    ap.add_argument("--no-plain", action="store_true", help="skip the 300-step plain workload (128x128 views)")
    ap.add_argument("--no-graph-compare", action="store_true",
                    help="skip timing replays of the captured step against launch-by-launch steps")
    ap.add_argument("--summarize-profile", nargs="+", metavar="DIR",
                    help="summarise rocprofv3 output directories of this benchmark (kernel trace with the "
                         "'timed_steps' range, --pmc runs) as JSON and exit; --warmup/--steps give the window")
    ap.add_argument("--calibration", nargs="+", metavar="DIR",
                    help="rocprofv3 --pmc output directories (FETCH_SIZE, WRITE_SIZE) of --copy-calibration")
    args = ap.parse_args()
    if args.summarize_profile:
        summarize_profile(args.summarize_profile, args.warmup, args.steps, args.calibration)
        return
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
        log("autotuned: " + str({k: tuning[k] for k in ("bwd_ppt", "fwd_batch", "bwd_batch", "depth_chunks", "ms")}))
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
    n_total = 0
    ctx.sync()
    written0 = slots_written(ctx)
    counts0 = frame_counts(ctx)
    rtx = _Roctx()
    rtx.push("timed_steps")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        tr.step()
        n_total += dg.n
    ctx.sync()
    elapsed = time.perf_counter() - t0
    rtx.pop()
    written = slots_written(ctx) - written0
    counts1 = frame_counts(ctx)
    m_total, v_total = counts1[0] - counts0[0], counts1[1] - counts0[1]
    timed = ctx.drain_stage_times()
    dtu_timed = min((first_timed_step + args.steps - 1) // 1000, 3)
    # stage breakdown over a separate window of steps that times every stage
    ctx.set_stage_mask(None)
    sw = {"M": 0, "V": 0, "N": 0, "B": 0}
    for _ in range(args.stage_steps):
        tr.step()
        f = ctx.frame()
        sw["M"] += f.num_isects
        sw["V"] += f.num_visible
        sw["N"] += dg.n
        sw["B"] += (min(tr.step_count // 1000, 3) + 1) ** 2
    stages = ctx.drain_stage_times()
    ctx.enable_stage_timing(False)
    loss = tr.step(read_loss=True)
    log(f"timed {args.steps} steps in {elapsed:.3f} s; gaussians {n_start} -> {dg.n}")

    K = args.steps
    iters = K / elapsed
    P = W_IMG * H_IMG
    T = ((W_IMG + 15) // 16) * ((H_IMG + 15) // 16)
    from msplat import autotune
    cur = autotune.current(ctx)
    # the dominant kernel: VALU-bound (per-pixel compositing work); its launch time from
    # the timed steps' events, its VALU instructions per launch from the counter summary
    dom_ms, dom_calls = timed[DOMINANT]
    avg_ms = dom_ms / max(dom_calls, 1)
    M, V, N = m_total / K, v_total / K, n_total / K
    hbm_8d = alg_bytes_8d(DOMINANT, M, V, N, P, T, 1)
    hbm_build = alg_bytes_build(DOMINANT, M, V, N, P, T, 1, written / max(dom_calls, 1))
    prof = profiled_counters(cur, args.warmup, args.steps)
    valu = prof.get("SQ_INSTS_VALU") if prof else None
    achieved = valu * 64 * 2 / (avg_ms * 1e-3) / 1e12 if valu else None
    roofline = {
        "bound": "valu", "kernel": bwd_kernel_name(cur["bwd_ppt"], cur["bwd_batch"]),
        "achieved": round(achieved, 3) if achieved else None, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
        "frac": round(achieved / FP32_PEAK_TFLOPS, 4) if achieved else None,
        "traffic": prof.get("hbm_bytes") if prof else None,
        "achieved_basis": "wave64 VALU instructions per launch (SQ_INSTS_VALU, counter summary of the same steps) "
                          "x 64 lanes x 2 (each counted as an FMA) / the timed steps' average launch time",
        "valu_insts_per_launch": valu,
        "valu_issue_util": round(valu * 4 / (SIMDS * CLOCK_HZ * avg_ms * 1e-3), 4) if valu else None,
        "valu_busy_pct": prof.get("VALUBusy_pct") if prof else None,
        "lds_util_pct": prof.get("LdsUtil_pct") if prof else None,
        "lds_bank_conflict_pct": prof.get("LDSBankConflict_pct") if prof else None,
        "lds_insts_per_launch": prof.get("SQ_INSTS_LDS") if prof else None,
        "counter_source": os.path.relpath(COUNTERS, ROOT) if prof else None,
        "counter_note": None if prof else "counters are profiled on --warmup 5 --steps 20 (steps 6-25); "
                                          "other windows report no VALU figures",
        "avg_launch_ms": round(avg_ms, 4),
        "intersections_per_launch": int(M), "visible_per_launch": int(V), "pixels": P,
        "slots_written_per_launch": int(written / max(dom_calls, 1)),
        "hbm": {"algorithmic_bytes_per_launch": int(hbm_8d), "achieved_gbs": round(hbm_8d / (avg_ms * 1e-3) / 1e9, 1),
                "frac": round(hbm_8d / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                "build_bytes_per_launch": int(hbm_build),
                "build_frac": round(hbm_build / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
        "mpix_per_s": round(P * iters / 1e6, 1),
        "sh_degree_in_use": dtu_timed,
    }
    # HBM rate of every stage over the stage window, by the per-unit formulas
    ns = max(args.stage_steps, 1)
    sM, sV, sN, sB = sw["M"] / ns, sw["V"] / ns, sw["N"] / ns, sw["B"] / ns
    stage_rows = {}
    for k, (ms, calls) in stages.items():
        if not calls:
            continue
        per = ms / calls
        b8, bb = alg_bytes_8d(k, sM, sV, sN, P, T, sB), alg_bytes_build(k, sM, sV, sN, P, T, sB)
        stage_rows[k] = {"ms": round(per, 4), "bytes": int(b8), "gbs": round(b8 / (per * 1e-3) / 1e9, 1),
                         "frac": round(b8 / (per * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        if bb != b8:
            stage_rows[k].update(build_bytes=int(bb), build_gbs=round(bb / (per * 1e-3) / 1e9, 1))
    roofline["stages"] = stage_rows
    roofline["stages_window"] = {"steps": args.stage_steps, "intersections": int(sM), "visible": int(sV),
                                 "gaussians": int(sN), "sh_bases": round(sB, 2)}
    heaviest = max(stages, key=lambda k: stages[k][0]) if stage_rows else DOMINANT
    if heaviest != DOMINANT:
        log(f"note: the stage breakdown's heaviest stage is {heaviest}, not {DOMINANT}")
    graph = None
    if not args.no_graph_compare:
        graph = time_step_graph(ctx, tr)
        log(f"captured step: {graph['graph_iter_per_s']} iter/s replayed, {graph['eager_iter_per_s']} launch by launch")
    sh3 = None
    if not args.no_sh3:
        sh3 = round(time_sh3_steps(ctx, dg, cams, targets), 2)
        log(f"SH degree 3 in use: {sh3} iter/s")
    batched = None
    if not args.no_batched:
        log("batched views: 1M gaussians, 16 views per launch")
        batched = bench_batched(args.batched_steps)
        log(f"batched: {batched['train_views_per_s']} train views/s, {batched['render_views_per_s']} render views/s")
    serving = None
    if not args.no_serving:
        import shutil
        import tempfile
        from msplat.scene_io import save_ply
        tmpdir = tempfile.mkdtemp(prefix="msplat_bench_")
        try:
            ply = os.path.join(tmpdir, "street_trained.ply")
            save_ply(ctx, dg, ply, step=tr.step_count)
            log("serving: 4 resident scenes, Poisson arrivals; then the trained street scene from its PLY file")
            serving = bench_serving(duration=args.serving_seconds, trained_ply=ply, trained_steps=tr.step_count)
        finally:
            shutil.rmtree(tmpdir, ignore_errors=True)
    scene_files = None
    if not args.no_scene_files:
        import shutil
        import tempfile
        tmpdir = tempfile.mkdtemp(prefix="msplat_files_")
        try:
            scene_files = bench_scene_files(ctx, dg, tr.step_count, tmpdir, cpu=not args.no_cpu_baseline)
            log("scene files: " + json.dumps(scene_files))
        finally:
            shutil.rmtree(tmpdir, ignore_errors=True)
    prior = None
    if not args.no_prior:
        log("prior-guided steps: street preset, SH degree 1")
        prior = time_prior_steps(ctx, dg, cams, targets)
        log(f"prior-guided: {prior['iter_per_s']} iter/s, plain SH1 {prior['plain_sh1_iter_per_s']} iter/s")
    images = None
    if not args.no_images:
        images = bench_images(cpu=not args.no_cpu_baseline)
        log("images: " + json.dumps(images["ops"]))
    plain = None
    if not args.no_plain:
        log("plain workload: 20k gaussians, 8 views at 128x128, 300 steps")
        plain = bench_plain()
        log(f"plain: {plain['per_tile']['iter_per_s']} iter/s per tile, "
            f"{plain['depth_chunked']['iter_per_s']} iter/s depth-chunked")
    cpu = None
    if not args.no_cpu_baseline:
        log("timing the CPU baseline")
        cpu = cpu_baseline_sample(dg, cams, targets, first_timed_step, args.cpu_seconds, plain)
        if prior:
            prior["cpu_reference"] = cpu_prior_sample(dg, cams, targets, prior)
    if prior:
        del prior["_cpu"]
    if plain:
        del plain["targets"]
    out = {
        "metric": METRIC, "value": round(iters, 3), "unit": "iter/s", "n_gpus": 1, "steps": K,
        "warmup": args.warmup, "ms_per_step": round(1e3 * elapsed / K, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "street-scale single-view training: 475,115 SH3 Gaussians, 1264x832, one view "
                               "per step, densify every 100 from step 500 (grad 2e-4, prune 0.005), 3000-step "
                               "schedule",
                   "gaussians_start": n_start, "gaussians_end": dg.n, "steps_timed": [first_timed_step,
                                                                                      first_timed_step + K - 1],
                   "sh_degree_in_use": dtu_timed, "mean_intersections_per_step": int(M),
                   "densify_events": tr.densify_log, "final_loss": loss, "setup_s": round(setup_s, 1),
                   "tuning": {k: tuning[k] for k in ("bwd_ppt", "fwd_batch", "bwd_batch", "depth_chunks")}
                   if tuning else None},
        "mpix_per_s": round(P * iters / 1e6, 1),
        "sh3_iter_per_s": sh3,
        "prior_step": prior,
        "images": images,
        "scene_files": scene_files,
        "step_graph": graph,
        "roofline": roofline,
        "batched": batched,
        "serving": serving,
        "plain": plain,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
