"""Depth-chunked rasterization for views with few tiles (MSPLAT_TUNE_CHUNKS; the chunked
form of msplat_metal.mm:782-907): each tile's list split into K parts, one workgroup per
(tile, part), merged per pixel. Against the one-workgroup-per-tile form and the oracle:
images, transmittance and every gradient within float rounding (1e-5 / 1e-4 relative),
last indices equal; opaque scenes (terminations inside chunks), more chunks than
records, and the per-frame choice."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BG = (0.61, 0.01, 0.4)


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def _scene(opaque: bool, n=20000, seed=1000):
    from msplat.scenes import random_scene, ring_cameras
    sc = random_scene(seed, n, degree=1, log_scale=(-3.6, 0.5))
    if opaque:
        sc["opac"] = (sc["opac"] + 3.0).astype(np.float32)
    return sc, ring_cameras(8, 4.0, 128, 128, 120.0, 120.0, dtu=1)[3]


def _frame(ctx, dg, cam, chunks, v_img):
    from msplat._abi import TUNE_CHUNKS
    from msplat.backend import render, render_backward
    ctx.set_tuning(TUNE_CHUNKS, chunks)
    render(ctx, dg, cam, BG)
    f = ctx.frame()
    H, W = cam["H"], cam["W"]
    out = dict(img=ctx.download(f.image, (H, W, 3), np.float32), T=ctx.download(f.final_T, (H, W), np.float32),
               idx=ctx.download(f.final_idx, (H, W), np.int32), chunks=f.depth_chunks, M=f.num_isects)
    out.update(render_backward(ctx, dg, cam, BG, v_img))
    ctx.set_tuning(TUNE_CHUNKS, 0)
    return out


@pytest.mark.parametrize("opaque", [False, True])
@pytest.mark.parametrize("chunks", [2, 7, 32])
def test_chunked_equals_per_tile(ctx, opaque, chunks):
    from msplat._abi import DeviceArray
    from msplat.backend import DeviceGaussians
    sc, cam = _scene(opaque)
    dg = DeviceGaussians(ctx, sc)
    v = DeviceArray.from_numpy(ctx, np.random.default_rng(5).normal(0, 1, (128, 128, 3)).astype(np.float32))
    one, many = _frame(ctx, dg, cam, 1, v), _frame(ctx, dg, cam, chunks, v)
    assert one["chunks"] == 1 and many["chunks"] == chunks and one["M"] > 64 * 64
    same_idx = np.mean(one["idx"] == many["idx"])
    errs = {k: _rel(many[k], one[k]) for k in ("img", "T", "means", "scales", "quats", "dc", "rest", "opac")}
    print(f"\nopaque={opaque} K={chunks}: last index equal on {same_idx:.6f} of pixels;",
          {k: f"{e:.1e}" for k, e in errs.items()})
    assert same_idx > 0.999
    assert errs["img"] < 1e-5 and errs["T"] < 1e-4
    for k in ("means", "scales", "quats", "dc", "rest", "opac"):
        assert errs[k] < 1e-4, (k, errs[k])


def test_chunked_train_step_matches_oracle(ctx, oracle):
    """A train step with 16 chunks per tile against the oracle's step on the same
    inputs (the parity bars of the one-view path)."""
    from msplat._abi import DeviceArray, TUNE_CHUNKS
    from msplat.backend import DeviceGaussians, train_step
    sc, cam = _scene(True, seed=1003)
    cam = dict(cam, dtu=1)
    dg = DeviceGaussians(ctx, sc)
    st = dg.to_numpy()
    gt = np.random.default_rng(6).uniform(0, 1, (128, 128, 3)).astype(np.float32)
    steps = np.array([1.6e-4, 5e-3, 1e-3, 2.5e-3, 1.25e-4, 5e-2], np.float32)
    bc2 = np.full(6, np.float32(np.sqrt(1 - 0.999)), np.float32)
    ctx.set_tuning(TUNE_CHUNKS, 16)
    try:
        loss = train_step(ctx, dg, cam, BG, DeviceArray.from_numpy(ctx, gt), steps, bc2)
        assert ctx.frame().depth_chunks == 16
    finally:
        ctx.set_tuning(TUNE_CHUNKS, 0)
    ref_loss, _, _ = oracle.train_step(st, cam, np.array(BG, np.float32), gt, steps, bc2)
    assert abs(loss - ref_loss) <= 1e-5 * abs(ref_loss)
    got = dg.to_numpy()
    for k in ("means", "scales", "quats", "dc", "rest", "opac"):
        for pre in ("", "m1_", "m2_"):
            e = _rel(got[pre + k], st[pre + k].reshape(got[pre + k].shape))
            assert e < 1e-4, (pre + k, e)


def test_more_chunks_than_records_and_the_per_frame_choice(ctx):
    """64 chunks over short lists (most chunks empty) equal the per-tile form; a 128x128
    view (64 tiles) is chunked by default, a 1264x832 view (4108 tiles) is not."""
    from msplat._abi import DeviceArray
    from msplat.backend import DeviceGaussians, render
    from msplat.scenes import random_scene, ring_cameras
    sc = random_scene(7, 300, degree=1, log_scale=(-2.5, 0.3))
    cam = ring_cameras(1, 4.0, 40, 36, 40.0, 40.0, dtu=1)[0]
    dg = DeviceGaussians(ctx, sc)
    v = DeviceArray.from_numpy(ctx, np.random.default_rng(8).normal(0, 1, (36, 40, 3)).astype(np.float32))
    one, many = _frame(ctx, dg, cam, 1, v), _frame(ctx, dg, cam, 64, v)
    assert np.array_equal(one["idx"], many["idx"]) and _rel(many["img"], one["img"]) < 1e-5
    for k in ("means", "opac"):
        assert _rel(many[k], one[k]) < 1e-4
    big, bcam = _scene(False, n=60000, seed=1004)
    dgb = DeviceGaussians(ctx, big)
    render(ctx, dgb, bcam, BG)
    f = ctx.frame()
    assert f.depth_chunks == (1 if f.num_isects < 32 * 64 else max(2, min(64, f.num_isects // (16 * 64))))
    wide = dict(ring_cameras(1, 4.0, 1264, 832, 900.0, 900.0, dtu=1)[0])
    render(ctx, dgb, wide, BG)
    assert ctx.frame().depth_chunks == 1
