"""The Python surface of the reference's `_core` module (python/bindings.cpp:431-778) on
the HIP backend: the GPU primitive hooks against NumPy, as the reference's own tests
check them (tests/test_msplat.py:429-597), and GaussianTrainer training an in-memory
dataset with the resolution schedule, densification, evaluation, export and resume."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,bits,hi", [(1, 32, 10), (5, 32, 4), (1023, 32, 2 ** 30), (1025, 32, 7),
                                       (70_000, 32, 2 ** 30), (5_000, 64, 2 ** 63), (2_049, 64, 3)])
def test_gpu_radix_sort_hook(n, bits, hi):
    from msplat import _core
    keys = np.random.default_rng(n + bits).integers(0, hi, n, dtype=np.uint64)
    k, order = _core._gpu_radix_sort(keys, bits)
    ref = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(k, keys[ref])
    np.testing.assert_array_equal(order, ref)


def test_gpu_knn_hook_against_brute_force():
    from msplat import _core
    pts = np.random.default_rng(7).uniform(-1, 1, (3000, 3)).astype(np.float32)
    d2 = ((pts[:, None, :].astype(np.float64) - pts[None]) ** 2).sum(-1)
    np.fill_diagonal(d2, np.inf)
    ref = np.sqrt(np.sort(d2, axis=1)[:, :3]).mean(axis=1)
    np.testing.assert_allclose(_core._gpu_knn3_mean_dist(pts), ref, rtol=1e-4, atol=1e-6)


def _area(src, dst):
    s = src / dst
    w = np.zeros((dst, src))
    for d in range(dst):
        a, b = d * s, (d + 1) * s
        for i in range(int(np.floor(a)), min(int(np.ceil(b)), src)):
            w[d, i] = min(i + 1, b) - max(i, a)
    return w / w.sum(1, keepdims=True)


@pytest.mark.parametrize("src,dst", [((48, 64), (24, 32)), ((50, 75), (17, 29)), ((40, 60), (40, 60)),
                                     ((10, 12), (25, 31))])
def test_gpu_resize_area_hook(src, dst):
    from msplat import _core
    rng = np.random.default_rng(src[0] * 1000 + dst[1])
    rgba = rng.integers(0, 256, (*src, 4), dtype=np.uint8)
    wy, wx = _area(src[0], dst[0]), _area(src[1], dst[1])
    got = _core._gpu_resize_area(rgba, dst[1], dst[0])
    assert got.shape == (*dst, 3) and got.dtype == np.float32
    np.testing.assert_allclose(got, np.einsum("yh,hwc,xw->yxc", wy, rgba[..., :3] / 255.0, wx), atol=1e-5)
    img = rng.random((*src, 3), dtype=np.float32)
    np.testing.assert_allclose(_core._gpu_resize_area(img, dst[1], dst[0]),
                               np.einsum("yh,hwc,xw->yxc", wy, img.astype(np.float64), wx), atol=1e-5)


def test_gpu_undistort_hook_crops_to_valid_pixels():
    """Pincushion distortion: the crop starts inside the image and every output pixel
    maps back inside the source (the alpha = 0 crop); values match a float64 remap."""
    from msplat import _core
    h, w = 90, 120
    ys, xs = np.mgrid[0:h, 0:w] / np.array([h, w])[:, None, None]
    img = np.stack([np.sin(6 * xs) * 0.5 + 0.5, np.cos(5 * ys) * 0.5 + 0.5, xs * ys], -1).astype(np.float32)
    k, d = [100.0, 102.0, 61.0, 44.0], [0.25, 0.05, 0.001, -0.002, 0.01]
    got, rx, ry = _core._gpu_undistort(img, k, d)
    assert rx > 0 and ry > 0
    oh, ow = got.shape[:2]
    assert rx + ow <= w and ry + oh <= h
    yy, xx = np.mgrid[0:oh, 0:ow].astype(np.float64)
    x, y = (xx + rx + 0.5 - k[2]) / k[0], (yy + ry + 0.5 - k[3]) / k[1]
    r2 = x * x + y * y
    rad = 1 + d[0] * r2 + d[1] * r2 ** 2 + d[4] * r2 ** 3
    sx = (x * rad + 2 * d[2] * x * y + d[3] * (r2 + 2 * x * x)) * k[0] + k[2] - 0.5
    sy = (y * rad + d[2] * (r2 + 2 * y * y) + 2 * d[3] * x * y) * k[1] + k[3] - 0.5
    assert sx.min() > -1 and sx.max() < w and sy.min() > -1 and sy.max() < h
    x0, y0 = np.floor(sx).astype(int), np.floor(sy).astype(int)
    ax, ay = (sx - x0)[..., None], (sy - y0)[..., None]
    xa, xb, ya, yb = (np.clip(v, 0, m - 1) for v, m in ((x0, w), (x0 + 1, w), (y0, h), (y0 + 1, h)))
    ref = (img[ya, xa] * (1 - ax) + img[ya, xb] * ax) * (1 - ay) + (img[yb, xa] * (1 - ax) + img[yb, xb] * ax) * ay
    np.testing.assert_allclose(got, ref, atol=1e-4)


def test_gpu_metrics_and_pack_hooks():
    from msplat import _core
    rng = np.random.default_rng(3)
    a = rng.random((37, 53, 3), dtype=np.float32)
    b = np.clip(a + rng.normal(0, 0.05, a.shape), 0, 1).astype(np.float32)
    psnr, ssim, l1 = _core._gpu_image_metrics(a, b)
    mse = np.mean((a.astype(np.float64) - b) ** 2)
    assert abs(psnr - 10 * np.log10(1 / mse)) < 1e-6 and abs(l1 - np.abs(a - b.astype(np.float64)).mean()) < 1e-7
    assert 0.5 < ssim < 1
    img = rng.uniform(-0.2, 1.2, (21, 17, 3)).astype(np.float32)
    rgba = _core._gpu_pack_rgba8(img)
    np.testing.assert_array_equal(rgba[..., :3], (np.clip(img, 0, 1) * 255).astype(np.uint8))
    assert (rgba[..., 3] == 255).all()


def _dataset(eval_mode=True):
    """Eight 96x72 views on a ring around a 3,000-Gaussian scene; targets rendered from a
    second seeded scene; the initial cloud is that scene's means with their colours."""
    from msplat import _core
    from msplat._abi import Context
    from msplat.backend import DeviceGaussians, render_numpy
    from msplat.camera import look_at
    from msplat.scenes import random_scene
    ctx = Context(0)
    gt = random_scene(81, 3000, degree=1, log_scale=(-3.5, 0.4))
    gdg = DeviceGaussians(ctx, gt)
    poses, intr, imgs = [], [], []
    from msplat.camera import make_camera
    for k in range(8):
        th = 2 * np.pi * k / 8
        pose = look_at((3.0 * np.cos(th), 0.4, 3.0 * np.sin(th)), (0, 0, 0))
        poses.append(pose)
        intr.append((96, 72, 80.0, 80.0, 48.0, 36.0))
        imgs.append(render_numpy(ctx, gdg, make_camera(pose, 96, 72, 80.0, 80.0, 48.0, 36.0, dtu=1), (0.2, 0.3, 0.4)))
    rng = np.random.default_rng(82)
    pts = (gt["means"] + rng.normal(0, 0.05, gt["means"].shape)).astype(np.float32)
    cols = rng.integers(0, 256, (3000, 3), dtype=np.uint8)
    ctx.close()
    return _core.Dataset.from_arrays(poses, intr, imgs, pts, cols, eval_mode=eval_mode, test_every=4)


def test_gaussian_trainer_trains_evaluates_exports_and_resumes(tmp_path):
    from msplat import _core
    from msplat.scene_io import load_ply
    ds = _dataset()
    assert ds.num_train == 6 and ds.num_test == 2 and ds.image(0).shape == (72, 96, 3)
    with pytest.raises(NotImplementedError):
        _core.Dataset(str(tmp_path))
    cfg = _core.TrainingConfig(iterations=120, sh_degree=1, sh_degree_interval=40, num_downscales=1,
                               resolution_schedule=60, refine_every=30, warmup_length=30, bg_color=[0.2, 0.3, 0.4])
    tr = _core.GaussianTrainer(ds, cfg)
    assert tr.resolution(0, 1) == (48, 36, 2) and tr.resolution(0, 60) == (96, 72, 1)
    before = tr.evaluate()["psnr"]
    seen = []
    tr.train(seen.append, callback_every=40)
    assert [s.iteration for s in seen] == [40, 80, 120] and tr.iteration == 120
    after = tr.evaluate()
    print(f"\ntest-view PSNR {before:.2f} -> {after['psnr']:.2f} dB, {tr.splat_count} gaussians")
    assert after["psnr"] > before + 1.0  # 120 steps: 21.4 -> 23.7 dB measured
    img, rgba = tr.render(1), tr.render_rgba8(1)
    np.testing.assert_array_equal(rgba[..., :3], (np.clip(img, 0, 1) * 255).astype(np.uint8))
    np.testing.assert_array_equal(tr.render_from_pose(ds.camera_pose(1), 1), img)
    tr.export_ply(str(tmp_path / "s.ply"))
    dg, _ = load_ply(tr.ctx, str(tmp_path / "s.ply"))
    np.testing.assert_array_equal(dg.to_numpy()["means"], tr._tr.dg.to_numpy()["means"])
    tr.save_checkpoint(str(tmp_path / "s.msplat"))
    state = tr._tr.dg.to_numpy()
    assert tr.load_checkpoint(str(tmp_path / "s.msplat")) == 120
    for k in ("means", "opac", "m1_means", "m2_dc"):
        np.testing.assert_array_equal(tr._tr.dg.to_numpy()[k], state[k])


def _prior_dataset():
    """Eight 96x72 views of a 3,000-Gaussian scene in front of a sky gradient (the
    targets: the scene composited over a colour that varies with the ray's elevation), and
    per view the priors a depth/sky estimator would give: the scene's expected depth over
    its alpha at 48x36 (0 where the alpha is low) and a sky mask where the alpha is below
    0.05. The initial cloud: the scene's means, jittered."""
    from msplat import _core
    from msplat._abi import Context
    from msplat.backend import DeviceGaussians
    from msplat.camera import look_at, make_camera
    from msplat.prior import DevicePrior, aux_outputs, render_prior
    from msplat.scenes import random_scene
    ctx = Context(0)
    gt = random_scene(83, 3000, degree=1, log_scale=(-3.5, 0.4))
    gt["means"] = (gt["means"] * np.float32(0.6)).astype(np.float32)
    gdg = DeviceGaussians(ctx, gt)
    poses, intr, imgs, priors = [], [], [], []
    for k in range(8):
        th = 2 * np.pi * k / 8
        pose = look_at((3.0 * np.cos(th), 0.4, 3.0 * np.sin(th)), (0, 0, 0))
        cam = make_camera(pose, 96, 72, 80.0, 80.0, 48.0, 36.0, dtu=1)
        img = render_prior(ctx, gdg, cam, (0.0, 0.0, 0.0), DevicePrior(ctx, dict(aux=True)))
        depth, fT = aux_outputs(ctx, 72, 96)
        alpha = 1.0 - fT
        rows = np.linspace(0.0, 1.0, 72, dtype=np.float32)[:, None, None]
        sky = np.broadcast_to(np.array([0.5, 0.6, 0.9], np.float32) * (1.2 - 0.4 * rows), (72, 96, 3))
        imgs.append((img + fT[..., None] * sky).astype(np.float32))
        poses.append(pose)
        intr.append((96, 72, 80.0, 80.0, 48.0, 36.0))
        d = np.where(alpha > 0.5, depth / np.maximum(alpha, 1e-6), 0.0).astype(np.float32)
        aux = np.zeros((36, 48, 4), np.uint8)
        aux[..., 0] = 200
        aux[..., 1] = np.where(alpha[1::2, 1::2] < 0.05, 255, 0)
        aux[..., 2] = 255
        priors.append(dict(depth=np.ascontiguousarray(d[1::2, 1::2]), aux=aux, has_depth=True, has_sky=True,
                           has_mask=False))
    rng = np.random.default_rng(84)
    pts = (gt["means"] + rng.normal(0, 0.05, gt["means"].shape)).astype(np.float32)
    cols = rng.integers(0, 256, (3000, 3), dtype=np.uint8)
    ctx.close()
    return _core.Dataset.from_arrays(poses, intr, imgs, pts, cols, eval_mode=True, test_every=4, priors=priors)


def test_gaussian_trainer_with_priors(tmp_path):
    """The street preset's prior-guided options (depth prior with a decaying weight, sky
    alpha and fill terms, learned sky, exposure, needle cap) through GaussianTrainer: the
    step runs the prior path, the sky is initialised from the sky-mask pixels and learned,
    the expected depth follows the prior, and the exported sky is the texture in 8 bits."""
    import zlib
    from msplat import _core
    ds = _prior_dataset()
    cfg = _core.TrainingConfig(iterations=150, sh_degree=1, sh_degree_interval=50, refine_every=50,
                               warmup_length=50, bg_color=[0.0, 0.0, 0.0], depth_weight=0.2,
                               depth_weight_final=0.05, sky_alpha_weight=0.05, fill_weight=0.01, learn_sky=True,
                               sky_width=64, sky_height=16, exposure_compensation=True, max_scale_ratio=10.0)
    assert cfg.priors_enabled()
    tr = _core.GaussianTrainer(ds, cfg)
    sky0 = tr._tr.prior_state.numpy("sky_tex").copy()
    assert sky0.shape == (16, 64, 3) and sky0.std() > 0.01  # initialised from the sky pixels
    before = tr.evaluate()["psnr"]
    losses = []
    for _ in range(150):
        tr.step()
        if tr.iteration % 10 == 0:
            losses.append(tr.prior_losses())
    after = tr.evaluate()["psnr"]
    print(f"\nprior-guided: test PSNR {before:.2f} -> {after:.2f} dB; prior terms first {losses[0]} last "
          f"{losses[-1]}")
    # the learned sky (initialised from the sky pixels) already explains much of each view:
    # measured 26.0 -> 26.3 dB over 150 steps, with the depth and sky terms more than halved
    assert after > before
    assert all(np.isfinite(v) for d in losses for v in d.values())
    assert losses[-1]["depth"] < 0.7 * losses[0]["depth"] and losses[-1]["sky"] < 0.7 * losses[0]["sky"]
    assert np.abs(tr._tr.prior_state.numpy("sky_tex") - sky0).max() > 1e-4  # the sky learned
    s = tr._tr.dg.to_numpy()["scales"]
    assert np.all(s.max(1) <= s.sum(1) - s.max(1) - s.min(1) + np.log(np.float32(10.0)) + 1e-5)
    depth, alpha = tr.render_depth(0)
    assert depth.shape == (72, 96) and alpha.shape == (72, 96)
    d0 = ds._train_priors[0]["depth"]
    m = (alpha[1::2, 1::2] > 0.9) & (d0 > 0)
    assert m.sum() > 50
    rel = np.abs(depth[1::2, 1::2][m] - d0[m]) / d0[m]
    print(f"expected depth vs prior: median relative error {np.median(rel):.3f}")
    assert np.median(rel) < 0.1
    tr.export_sky(str(tmp_path / "sky.png"))
    data = open(tmp_path / "sky.png", "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    i = data.index(b"IDAT")
    n = int.from_bytes(data[i - 4:i], "big")
    raw = np.frombuffer(zlib.decompress(data[i + 4:i + 4 + n]), np.uint8).reshape(16, 1 + 64 * 3)
    want = (np.clip(tr._tr.prior_state.numpy("sky_tex") * np.float32(255), 0, 255) + np.float32(0.5)).astype(np.uint8)
    np.testing.assert_array_equal(raw[:, 1:].reshape(16, 64, 3), want)
