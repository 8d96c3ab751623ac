"""Autotuner for the rasterizer's kernel variants (msplat_hip_set_tuning).

The tunables are the backward rasterizer's pixels per thread (the work tile of one lane),
the forward and backward LDS batch depths (gaussian records staged per batch) and the
depth chunks per tile (the work unit of one workgroup along a tile's list: 0 = chosen per
frame, 1 = the whole list, 2 or 4 chunks). The binning tile stays 16x16: it decides which
gaussians reach a pixel. Every variant computes the same result (images and gradients
bit for bit for the batch depths; the pixels-per-thread variants within float rounding
of the per-wave sums, the chunk counts within float rounding of the transmittance
products), so the choice only moves time. The tuner times the caller's own step (a train step, a
render, a batched step) on the context's stream: coordinate descent over the tunables,
each tried at every value with the others at their best so far, a few timed repeats per
trial after a warm-up, the fastest kept. A result can be cached in a JSON file keyed by
device and workload.
"""
from __future__ import annotations

import json
import os
import time

from ._abi import TUNE_BWD_BATCH, TUNE_BWD_PPT, TUNE_CHUNKS, TUNE_FWD_BATCH, Context

SPACE = {
    "bwd_ppt": (TUNE_BWD_PPT, (4, 2, 1)),
    "fwd_batch": (TUNE_FWD_BATCH, (256, 128, 64)),
    "bwd_batch": (TUNE_BWD_BATCH, (64, 128, 32)),
    "depth_chunks": (TUNE_CHUNKS, (0, 1, 2, 4)),
}


def current(ctx: Context) -> dict:
    return {name: ctx.get_tuning(key) for name, (key, _) in SPACE.items()}


def apply(ctx: Context, choice: dict):
    for name, value in choice.items():
        if name in SPACE:
            ctx.set_tuning(SPACE[name][0], int(value))


def _time(ctx: Context, step, warmup: int, repeats: int, reset=None) -> float:
    if reset is not None:
        reset()
    for _ in range(warmup):
        step()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(repeats):
        step()
    ctx.sync()
    return (time.perf_counter() - t0) / repeats * 1e3


def autotune(ctx: Context, step, warmup: int = 2, repeats: int = 8, space: dict | None = None,
             cache: str | None = None, key: str | None = None, rounds: int = 1, min_gain: float = 0.02,
             reset=None) -> dict:
    """Picks the fastest variant for `step()` (a callable that enqueues one unit of the
    workload on ctx) and leaves it set on ctx. A variant replaces the best so far only
    when it is faster by more than `min_gain` (a fraction), so timing noise does not flip
    the choice. `reset()`, when given, runs before every trial: a step that changes its
    own input (a train step moves the parameters, and a scene early in training gets
    cheaper step by step) must start each trial from the same state, or the variants
    tried last look fastest. Returns {tunable: value, ..., "ms": best ms per step, "trials":
    [(choice, ms), ...], "cached": bool}."""
    space = space or SPACE
    if cache and key:
        try:
            with open(cache) as f:
                hit = json.load(f).get(key)
        except (OSError, ValueError):
            hit = None
        if hit:
            apply(ctx, hit)
            return dict(hit, cached=True)
    best = current(ctx)
    # one discarded trial first: the first steps on a context also grow its buffers and
    # load the kernels, and timing them would make every later variant look faster
    _time(ctx, step, warmup, repeats, reset)
    best_ms = _time(ctx, step, warmup, repeats, reset)
    trials = [(dict(best), best_ms)]
    for _ in range(rounds):
        for name, (tkey, values) in space.items():
            for v in values:
                if v == best[name]:
                    continue
                trial = dict(best, **{name: v})
                apply(ctx, trial)
                ms = _time(ctx, step, warmup, repeats, reset)
                trials.append((trial, ms))
                if ms < best_ms * (1.0 - min_gain):
                    best, best_ms = trial, ms
            apply(ctx, best)
    out = dict(best, ms=round(best_ms, 4), trials=[(c, round(m, 4)) for c, m in trials], cached=False)
    if cache and key:
        try:
            data = {}
            if os.path.exists(cache):
                with open(cache) as f:
                    data = json.load(f)
            data[key] = dict(best, ms=round(best_ms, 4))
            os.makedirs(os.path.dirname(os.path.abspath(cache)), exist_ok=True)
            with open(cache, "w") as f:
                json.dump(data, f, indent=1)
        except (OSError, ValueError):
            pass
    return out
