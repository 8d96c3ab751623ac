"""Tile culling in the rasterizers (MSPLAT_TUNE_CULL) changes no output bit: a scene
whose opacities reach below and around the 1/255 cut-off, with anisotropic splats,
renders the same image, final T and contributor indices, and gives the same gradients,
with culling on and off, for every backward pixels-per-thread variant and for a batch
of views. The CPU test checks that the scene does exercise the culling: a good share
of its (tile, gaussian) pairs have their alpha >= 1/255 ellipse outside the tile."""
import numpy as np
import pytest

BG = (0.3, 0.1, 0.6)
W, H = 157, 93


def _scene():
    from msplat.camera import look_at, make_camera
    from msplat.scenes import random_scene
    sc = random_scene(41, 8000, degree=3, log_scale=(-3.3, 0.7), opac=(-2.0, 2.5))
    cam = make_camera(look_at((2.3, 0.8, 2.7), (0.0, 0.0, 0.0)), W, H, 120.0, 118.0, dtu=3)
    return sc, cam


def test_cull_scene_has_pairs_outside_their_tiles(oracle):
    sc, cam = _scene()
    p = oracle.project(sc, cam)
    b = oracle.bin_sort(p, sc["opac"], cam)
    ts, te = b["tile_start"].astype(np.int64), b["tile_end"].astype(np.int64)
    tile = np.repeat(np.arange(ts.size), np.maximum(te - ts, 0))
    x, y, o = [b["pk_xyo"][:, k].astype(np.float64) for k in range(3)]
    a, bb, c = [b["pk_conic"][:, k].astype(np.float64) for k in range(3)]
    det = a * c - bb * bb
    lvl = np.minimum(5.55, np.log(np.maximum(255.0 * o, 1e-300)))
    ok = (det > 0) & (lvl > 0)
    hx = np.sqrt(np.where(ok, 2 * lvl * c / np.where(ok, det, 1), 0))
    hy = np.sqrt(np.where(ok, 2 * lvl * a / np.where(ok, det, 1), 0))
    x0 = 16.0 * (tile % cam["tiles_x"])
    y0 = 16.0 * (tile // cam["tiles_x"])
    miss = ~ok | (x + hx < x0) | (x - hx > x0 + 15) | (y + hy < y0) | (y - hy > y0 + 15)
    assert b["M"] > 20000 and miss.mean() > 0.15, (b["M"], miss.mean())
    assert (o < 1 / 255).sum() > 50  # some splats never pass the alpha cut-off


def _frame(ctx, dg, cam, v_img):
    from msplat.backend import render, render_backward
    ptr = render(ctx, dg, cam, BG)
    f = ctx.frame()
    out = [ctx.download(ptr, (H, W, 3), np.float32), ctx.download(f.final_T, (H, W), np.float32),
           ctx.download(f.final_idx, (H, W), np.int32)]
    return out, render_backward(ctx, dg, cam, BG, v_img)


@pytest.mark.gpu
def test_cull_changes_no_bit(ctx):
    from msplat._abi import TUNE_BWD_PPT, TUNE_CULL, DeviceArray
    from msplat.backend import DeviceGaussians
    sc, cam = _scene()
    dg = DeviceGaussians(ctx, sc)
    v_img = DeviceArray.from_numpy(ctx, np.random.default_rng(2).normal(0, 1e-3, (H, W, 3)).astype(np.float32))
    saved = (ctx.get_tuning(TUNE_BWD_PPT), ctx.get_tuning(TUNE_CULL))
    try:
        for ppt in (1, 2, 4):
            ctx.set_tuning(TUNE_BWD_PPT, ppt)
            ctx.set_tuning(TUNE_CULL, 0)
            ref_px, ref_gr = _frame(ctx, dg, cam, v_img)
            ctx.set_tuning(TUNE_CULL, 1)
            px, gr = _frame(ctx, dg, cam, v_img)
            assert ref_px[0].std() > 0.02
            for a, b in zip(px, ref_px):
                np.testing.assert_array_equal(a, b, err_msg=f"ppt {ppt}")
            for k in ref_gr:
                np.testing.assert_array_equal(gr[k], ref_gr[k], err_msg=f"{k} ppt {ppt}")
    finally:
        ctx.set_tuning(TUNE_BWD_PPT, saved[0])
        ctx.set_tuning(TUNE_CULL, saved[1])


@pytest.mark.gpu
def test_cull_changes_no_bit_batched(ctx):
    from msplat._abi import TUNE_CULL, DeviceArray
    from msplat.backend import BatchViews, DeviceGaussians, render_backward_batch, render_batch_numpy
    from msplat.camera import look_at, make_camera
    sc, cam = _scene()
    dg = DeviceGaussians(ctx, sc)
    cams = [cam, make_camera(look_at((-2.0, 0.4, 2.9), (0.0, 0.1, 0.0)), 90, 131, 100.0, 100.0, dtu=3)]
    bv = BatchViews(cams, 3, [BG])
    pix = sum(c["W"] * c["H"] for c in cams)
    v = DeviceArray.from_numpy(ctx, np.random.default_rng(4).normal(0, 1e-3, (pix, 3)).astype(np.float32))
    saved = ctx.get_tuning(TUNE_CULL)
    try:
        res = []
        for cull in (0, 1):
            ctx.set_tuning(TUNE_CULL, cull)
            imgs = render_batch_numpy(ctx, dg, bv)
            res.append((imgs, render_backward_batch(ctx, dg, bv, v)))
        for a, b in zip(res[0][0], res[1][0]):
            np.testing.assert_array_equal(a, b)
        for k in res[0][1]:
            np.testing.assert_array_equal(res[0][1][k], res[1][1][k], err_msg=k)
    finally:
        ctx.set_tuning(TUNE_CULL, saved)


@pytest.mark.gpu
def test_repeated_backward_of_one_frame(ctx):
    """A slot left zero by one backward is flagged zero again by the next backward of the
    same frame: two backward passes of one render, with cotangents zero in different
    halves of the image, give the gradients of a fresh render and backward each."""
    from msplat._abi import DeviceArray
    from msplat.backend import DeviceGaussians, render, render_backward
    sc, cam = _scene()
    dg = DeviceGaussians(ctx, sc)
    rng = np.random.default_rng(5)
    va = rng.normal(0, 1e-3, (H, W, 3)).astype(np.float32)
    vb = va.copy()
    va[:, W // 2:] = 0.0  # the first backward sees the left half only, the second the right half
    vb[:, : W // 2] = 0.0
    da, db = DeviceArray.from_numpy(ctx, va), DeviceArray.from_numpy(ctx, vb)
    fresh = []
    for d in (da, db):
        render(ctx, dg, cam, BG)
        fresh.append(render_backward(ctx, dg, cam, BG, d))
    render(ctx, dg, cam, BG)
    again = [render_backward(ctx, dg, cam, BG, d) for d in (da, db, da)]
    for got, want in zip(again, fresh + fresh[:1]):
        for k in want:
            assert np.abs(want[k]).max() > 0, k
            np.testing.assert_array_equal(got[k], want[k], err_msg=k)
