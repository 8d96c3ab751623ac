"""Kernel variants of the rasterizer give the same frame: every forward LDS batch depth
renders the same image, final T and contributor indices bit for bit; every backward
batch depth gives the same gradients bit for bit (per pixels-per-thread variant), and the
autotuner leaves a valid variant set on the context."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
BG = (0.3, 0.1, 0.6)


def _frame(ctx, dg, cam, v_img):
    from msplat._abi import DeviceArray
    from msplat.backend import render, render_backward
    p = render(ctx, dg, cam, BG)
    f = ctx.frame()
    H, W = cam["H"], cam["W"]
    img = ctx.download(p, (H, W, 3), np.float32)
    fT = ctx.download(f.final_T, (H, W), np.float32)
    fi = ctx.download(f.final_idx, (H, W), np.int32)
    grads = render_backward(ctx, dg, cam, BG, v_img)
    return img, fT, fi, grads


def test_variants_give_the_same_frame(ctx):
    from msplat._abi import TUNE_BWD_BATCH, TUNE_BWD_PPT, TUNE_FWD_BATCH, DeviceArray
    from msplat.backend import DeviceGaussians
    from msplat.camera import look_at, make_camera
    from msplat.scenes import random_scene
    sc = random_scene(31, 6000, degree=3, log_scale=(-3.4, 0.5))
    cam = make_camera(look_at((2.4, 0.9, 2.6), (0.0, 0.0, 0.0)), 131, 77, 95.0, 97.0, dtu=3)
    dg = DeviceGaussians(ctx, sc)
    v_img = DeviceArray.from_numpy(ctx, (np.random.default_rng(3).normal(0, 1e-3, (77, 131, 3))).astype(np.float32))
    saved = [ctx.get_tuning(k) for k in (TUNE_BWD_PPT, TUNE_FWD_BATCH, TUNE_BWD_BATCH)]
    try:
        for ppt in (1, 2, 4):
            ref = None
            for fb in (64, 128, 256):
                for bb in (32, 64, 128):
                    ctx.set_tuning(TUNE_BWD_PPT, ppt)
                    ctx.set_tuning(TUNE_FWD_BATCH, fb)
                    ctx.set_tuning(TUNE_BWD_BATCH, bb)
                    got = _frame(ctx, dg, cam, v_img)
                    if ref is None:
                        ref = got
                        continue
                    for a, b in zip(got[:3], ref[:3]):
                        np.testing.assert_array_equal(a, b, err_msg=f"ppt {ppt} fwd {fb} bwd {bb}")
                    for k in ref[3]:
                        np.testing.assert_array_equal(got[3][k], ref[3][k], err_msg=f"{k} ppt {ppt} fb {fb} bb {bb}")
    finally:
        for k, v in zip((TUNE_BWD_PPT, TUNE_FWD_BATCH, TUNE_BWD_BATCH), saved):
            ctx.set_tuning(k, v)


def test_packed_sort_keys_give_the_same_frame(ctx):
    """The tile sort on packed (tile << id bits | id) keys and on tile keys plus id values
    give the same sorted lists, tile ranges, image and gradients, bit for bit, for a
    scene whose ids need 13 bits (packed) and for the same frame with packing off."""
    from msplat._abi import FRAME_KEYS64, TUNE_PACK_KEYS, DeviceArray
    from msplat.backend import DeviceGaussians
    from msplat.camera import look_at, make_camera
    from msplat.scenes import random_scene
    sc = random_scene(35, 6000, degree=3, log_scale=(-3.4, 0.5))
    cam = make_camera(look_at((2.4, 0.9, 2.6), (0.0, 0.0, 0.0)), 131, 77, 95.0, 97.0, dtu=3)
    dg = DeviceGaussians(ctx, sc)
    v_img = DeviceArray.from_numpy(ctx, (np.random.default_rng(4).normal(0, 1e-3, (77, 131, 3))).astype(np.float32))
    saved = ctx.get_tuning(TUNE_PACK_KEYS)
    out = []
    try:
        ctx.set_flags(FRAME_KEYS64)
        for pack in (1, 0):
            ctx.set_tuning(TUNE_PACK_KEYS, pack)
            img, fT, fi, grads = _frame(ctx, dg, cam, v_img)
            f = ctx.frame()
            M, T = f.num_isects, f.num_tiles
            lists = (ctx.download(f.isect_gid, (M,), np.int32), ctx.download(f.isect_keys, (M,), np.uint64),
                     ctx.download(f.tile_start, (T + 1,), np.int32))
            out.append((img, fT, fi, grads, lists))
    finally:
        ctx.set_flags(0)
        ctx.set_tuning(TUNE_PACK_KEYS, saved)
    (a, b) = out
    assert a[4][0].size > 10_000
    for x, y in zip(a[4], b[4]):
        np.testing.assert_array_equal(x, y)
    for x, y in zip(a[:3], b[:3]):
        np.testing.assert_array_equal(x, y)
    for k in a[3]:
        np.testing.assert_array_equal(a[3][k], b[3][k], err_msg=k)


def test_packed_sort_keys_at_street_scale(ctx):
    """475,115 Gaussians (19 id bits) on 1264x832 (4,108 tiles, 13 bits): the packed key
    uses all 32 bits. Packed and unpacked sorts give the same frame, bit for bit."""
    from msplat._abi import TUNE_PACK_KEYS, DeviceArray
    from msplat.backend import DeviceGaussians
    from msplat.scenes import street_cameras, street_points
    n, W, H = 475_115, 1264, 832
    assert (n - 1) >> 18 == 1 and ((W + 15) // 16) * ((H + 15) // 16) > 4096
    pts, rgb = street_points(0, n)
    dg = DeviceGaussians.from_points(ctx, pts, rgb, degree=3)
    cam = dict(street_cameras(4, W, H, 900.0, 900.0)[1], dtu=3)
    v_img = DeviceArray.from_numpy(ctx, (np.random.default_rng(5).normal(0, 1e-6, (H, W, 3))).astype(np.float32))
    saved = ctx.get_tuning(TUNE_PACK_KEYS)
    out = []
    try:
        for pack in (1, 0):
            ctx.set_tuning(TUNE_PACK_KEYS, pack)
            out.append(_frame(ctx, dg, cam, v_img) + (ctx.frame().num_isects,))
    finally:
        ctx.set_tuning(TUNE_PACK_KEYS, saved)
    a, b = out
    assert a[4] == b[4] and a[4] > 1_000_000, (a[4], b[4])
    for x, y in zip(a[:3], b[:3]):
        np.testing.assert_array_equal(x, y)
    for k in a[3]:
        np.testing.assert_array_equal(a[3][k], b[3][k], err_msg=k)


def test_autotuner_picks_a_valid_variant(ctx, tmp_path):
    from msplat import autotune
    from msplat.backend import DeviceGaussians, render
    from msplat.camera import look_at, make_camera
    from msplat.scenes import random_scene
    sc = random_scene(32, 20000, degree=3, log_scale=(-3.6, 0.5))
    cam = make_camera(look_at((2.4, 0.9, 2.6), (0.0, 0.0, 0.0)), 320, 240, 250.0, 250.0, dtu=3)
    dg = DeviceGaussians(ctx, sc)
    saved = autotune.current(ctx)
    try:
        cache = str(tmp_path / "tune.json")
        res = autotune.autotune(ctx, lambda: render(ctx, dg, cam, BG), warmup=1, repeats=3, cache=cache, key="k")
        assert not res["cached"] and len(res["trials"]) == 10
        for name, (_, values) in autotune.SPACE.items():
            assert res[name] in values and autotune.current(ctx)[name] == res[name]
        again = autotune.autotune(ctx, lambda: render(ctx, dg, cam, BG), cache=cache, key="k")
        assert again["cached"] and all(again[n] == res[n] for n in autotune.SPACE)
    finally:
        autotune.apply(ctx, saved)


def test_stage_mask_times_only_the_named_stages(ctx):
    from msplat._abi import STAGE_NAMES, DeviceArray
    from msplat.backend import DeviceGaussians, render, render_backward
    from msplat.camera import look_at, make_camera
    from msplat.scenes import random_scene
    sc = random_scene(33, 3000, degree=3, log_scale=(-3.4, 0.5))
    cam = make_camera(look_at((2.4, 0.9, 2.6), (0.0, 0.0, 0.0)), 96, 64, 80.0, 80.0, dtu=3)
    dg = DeviceGaussians(ctx, sc)
    v_img = DeviceArray.from_numpy(ctx, np.full((64, 96, 3), 1e-3, np.float32))
    try:
        ctx.enable_stage_timing(True)
        for stages in (None, ("rast_bwd",), ("prefix_sort_pack", "rast_fwd")):
            ctx.set_stage_mask(stages)
            ctx.drain_stage_times()
            for _ in range(3):
                render(ctx, dg, cam, BG)
                render_backward(ctx, dg, cam, BG, v_img)
            st = ctx.drain_stage_times()
            want = set(STAGE_NAMES if stages is None else stages) - {"loss_fwd_bwd"}  # no loss stage here
            for name, (ms, calls) in st.items():
                if name in want:
                    assert calls == 3 and ms > 0, (stages, name, st[name])
                else:
                    assert calls == 0 and ms == 0, (stages, name, st[name])
        with pytest.raises(RuntimeError):
            from msplat._abi import check, lib
            check(lib().msplat_hip_set_stage_mask(ctx.ptr, 1 << len(STAGE_NAMES)))
    finally:
        ctx.set_stage_mask(None)
        ctx.enable_stage_timing(False)
