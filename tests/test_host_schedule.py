"""Host-side math of the reference's Model: the camera setup (Model::prepareCam,
model.cpp:498-546), the means learning-rate decay (Model::schedulersStep,
model.cpp:140-143), the Adam bias corrections (model.cpp:638-651) and the refine cadence
(Model::afterTrain, model.cpp:199-245).

The reference's Model does not build here (model.hpp includes input_data.hpp, which
needs nlohmann/json), so the oracle restates these functions in C++ and this file pins
them by what the mathematics fixes: a pinhole projection for the matrices, the end
points and geometric midpoint of the log-linear decay, the double-rounded bias
correction at t = 1, and the enumerated refine steps of two schedules. The product's
Python (msplat.camera, msplat.trainer, msplat.backend) is then checked against the
oracle bit for bit.
"""
import numpy as np
import pytest

F = np.float32


def _random_pose(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                  [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                  [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
    m = np.eye(4)
    m[:3, :3], m[:3, 3] = R, rng.uniform(-20, 20, 3)
    return m.astype(np.float32)


def test_prepare_cam_axis_aligned_closed_form(oracle):
    """An OpenGL camera at t looking down -z with +y up: the OpenCV view matrix is
    diag(1, -1, -1) with translation (-tx, ty, tz), exactly."""
    t = np.array([1.5, -2.25, 7.0], F)
    c2w = np.eye(4, dtype=F)
    c2w[:3, 3] = t
    r = oracle.prepare_cam(c2w, 640, 480, 500.0, 510.0, 320.0, 240.0)
    vm = r["viewmat"].reshape(4, 4)
    np.testing.assert_array_equal(vm, np.array([[1, 0, 0, -t[0]], [0, -1, 0, t[1]], [0, 0, -1, t[2]], [0, 0, 0, 1]],
                                               F))
    np.testing.assert_array_equal(r["cam_pos"], t)
    pm = r["projmat"].reshape(4, 4)
    # P = diag(2fx/W, 2fy/H, (f+n)/(f-n)) with w = z_cam: row 0 is (2fx/W) * viewmat row 0
    np.testing.assert_allclose(pm[0], 2 * 500.0 / 640 * vm[0], rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(pm[1], 2 * 510.0 / 480 * vm[1], rtol=1e-6, atol=1e-6)
    np.testing.assert_array_equal(pm[3], vm[2])
    np.testing.assert_allclose(pm[2], (1000 + 1e-3) / (1000 - 1e-3) * vm[2] + np.array([0, 0, 0, -1.0 / (1000 - 1e-3)]),
                               rtol=1e-5, atol=1e-7)


def test_prepare_cam_is_a_pinhole(oracle):
    """For any pose, the projection matrix maps a point to the pinhole pixel
    fx x/z + cx (through the reference's ndc2pix, cpu:140) and the view matrix inverts
    the flipped camera-to-world."""
    rng = np.random.default_rng(0)
    flip = np.diag([1.0, -1.0, -1.0, 1.0])
    for _ in range(50):
        c2w = _random_pose(rng)
        W, H = int(rng.integers(16, 4000)), int(rng.integers(16, 3000))
        fx, fy = rng.uniform(100, 3000, 2)
        cx, cy = W / 2 + rng.uniform(-9, 9), H / 2 + rng.uniform(-9, 9)
        r = oracle.prepare_cam(c2w, W, H, fx, fy, cx, cy)
        vm, pm = r["viewmat"].reshape(4, 4).astype(np.float64), r["projmat"].reshape(4, 4).astype(np.float64)
        np.testing.assert_allclose(vm @ (c2w.astype(np.float64) @ flip), np.eye(4), atol=2e-6)
        p = np.r_[c2w[:3, 3] + rng.normal(0, 5, 3) + 30 * (c2w[:3, :3] @ [0, 0, -1]), 1.0]
        xc = vm @ p
        assert xc[2] > 0
        h = pm @ p
        px = 0.5 * W * (h[0] / h[3]) + cx - 0.5
        py = 0.5 * H * (h[1] / h[3]) + cy - 0.5
        np.testing.assert_allclose([px, py], [fx * xc[0] / xc[2] + cx - 0.5, fy * xc[1] / xc[2] + cy - 0.5],
                                   rtol=1e-5, atol=1e-3 * max(W, H) / 1000)
        np.testing.assert_array_equal(r["cam_pos"], c2w[:3, 3])


def test_prepare_cam_pyramid_levels(oracle):
    """Downscale factor sf: levelSize = floor(W / sf), intrinsics scaled by the exact
    per-axis ratio (1/4 for 1264 / 4; 632/1265 for an odd width)."""
    r = oracle.prepare_cam(np.eye(4, dtype=F), 1264, 832, 900.0, 900.0, 632.0, 416.0, sf=4)
    assert (r["W"], r["H"]) == (316, 208)
    assert (r["fx"], r["fy"], r["cx"], r["cy"]) == (225.0, 225.0, 158.0, 104.0)
    r = oracle.prepare_cam(np.eye(4, dtype=F), 1265, 833, 900.0, 901.0, 632.5, 416.5, sf=2)
    assert (r["W"], r["H"]) == (632, 416)
    assert r["fx"] == F(F(900.0) * F(F(632) / F(1265)))
    assert r["fy"] == F(F(901.0) * F(F(416) / F(833)))
    r1 = oracle.prepare_cam(np.eye(4, dtype=F), 1265, 833, 900.0, 901.0, 632.5, 416.5, sf=1)
    assert (r1["W"], r1["H"], r1["fx"]) == (1265, 833, 900.0)


def test_means_lr_closed_forms(oracle):
    """exp(lerp(log a, log b, t)): a at t = 0, b at t >= 1, sqrt(ab) at t = 1/2, and
    monotone in between."""
    a, b = 1.6e-4, 1.6e-6
    # exp(log a) in float: log a ~ -8.7 carries half an ulp of 2^3, i.e. ~5e-7 relative
    assert abs(oracle.means_lr(0, 30000) - a) <= 1e-6 * a
    assert abs(oracle.means_lr(30000, 30000) - b) <= 1e-6 * b
    assert oracle.means_lr(45000, 30000) == oracle.means_lr(30000, 30000)
    assert abs(oracle.means_lr(15000, 30000) - np.sqrt(a * b)) <= 1e-6 * np.sqrt(a * b)
    lrs = np.array([oracle.means_lr(s, 3000) for s in range(3001)])
    assert np.all(np.diff(lrs) <= 0) and lrs[0] > lrs[-1]


def test_adam_scalars_closed_forms(oracle):
    """bc = 1 - beta^t with the float beta promoted to double: at t = 1, 1 - 0.9f is
    0.10000002384185791 exactly (a float), so the step is lr / that; for large t the
    corrections vanish and the step is lr."""
    s, b = oracle.adam_scalars(F(5e-3), 1)
    assert s == F(F(5e-3) / F(0.10000002384185791))
    assert b == F(np.sqrt(F(1.0 - float(F(0.999)))))
    s, b = oracle.adam_scalars(F(5e-3), 100000)
    assert (s, b) == (F(5e-3), F(1.0))
    # the double power, not a float one: at some t the two round to different floats
    diff = [t for t in range(1, 3000) if F(1.0 - float(F(0.999)) ** t) != F(1) - F(0.999) ** F(t)]
    assert diff
    for t in diff[:20]:
        _, b = oracle.adam_scalars(F(1e-3), t)
        assert b == F(np.sqrt(F(1.0 - float(F(0.999)) ** t)))


def test_refine_cadence_enumerated(oracle):
    """The street schedule (3000 steps, refine 100, warmup 499, 64 views): densify at
    500..1400 (stop_split_at = 1500), no opacity reset. The default schedule (30000
    steps, warmup 500, 64 views): resets at 3100, 6100, 9100, 12100; densify at refine
    steps in (500, 15000) whose remainder mod 3000 exceeds 164."""
    def flags(step, warmup, stop):
        return oracle.refine_flags(step, 100, warmup, 30, stop, 64, 4000)

    street = [s for s in range(1, 3001) if flags(s, 499, 1500) & 1]
    assert street == list(range(500, 1500, 100))
    assert not any(flags(s, 499, 1500) & 2 for s in range(1, 3001))
    assert all(flags(s, 499, 1500) & 4 for s in range(500, 3001, 100))
    resets = [s for s in range(1, 30001) if flags(s, 500, 15000) & 2]
    assert resets == [3100, 6100, 9100, 12100]
    dens = [s for s in range(1, 30001) if flags(s, 500, 15000) & 1]
    assert dens == [s for s in range(600, 15000, 100) if s % 3000 > 164]
    huge = [s for s in dens if flags(s, 500, 15000) & 16]
    assert huge == [s for s in dens if s > 3000]
    screen = [s for s in dens if flags(s, 500, 15000) & 8]
    assert screen == [s for s in dens if s < 4000]


# ── the product's host code against the (pinned) oracle, bit for bit ──

def test_camera_matrices_equal_oracle(oracle):
    from msplat.camera import make_camera
    rng = np.random.default_rng(1)
    for i in range(300):
        c2w = _random_pose(rng)
        W, H = int(rng.integers(1, 4000)), int(rng.integers(1, 3000))
        fx, fy = (float(F(v)) for v in rng.uniform(50, 3000, 2))
        cx, cy = float(F(W / 2 + rng.uniform(-9, 9))), float(F(H / 2 + rng.uniform(-9, 9)))
        sf = int(rng.choice([1, 1, 2, 4, 8]))
        got = make_camera(c2w, W, H, fx, fy, cx, cy, downscale=sf)
        ref = oracle.prepare_cam(c2w, W, H, fx, fy, cx, cy, sf=sf)
        for k in ("viewmat", "projmat", "cam_pos"):
            np.testing.assert_array_equal(np.asarray(got[k], F).ravel(), ref[k], err_msg=f"{k} case {i}")
        for k in ("fx", "fy", "cx", "cy", "W", "H"):
            assert got[k] == ref[k], (k, i, got[k], ref[k])


@pytest.mark.parametrize("max_steps", [300, 3000, 30000])
def test_means_lr_equals_oracle(oracle, max_steps):
    from msplat.trainer import means_lr
    for s in range(0, max_steps + 2):
        assert F(means_lr(s, max_steps)) == oracle.means_lr(s, max_steps), s


def test_adam_coeffs_equal_oracle(oracle):
    from msplat.backend import adam_coeffs
    from msplat.trainer import LR_INIT
    for t in list(range(1, 5000)) + list(range(5000, 30001, 7)):
        steps, bc2 = adam_coeffs(LR_INIT, t)
        for i, lr in enumerate(LR_INIT):
            s, b = oracle.adam_scalars(F(lr), t)
            assert steps[i] == s and bc2[i] == b, (t, i)


def test_refine_flags_equal_oracle(oracle):
    from msplat.trainer import TrainConfig, refine_flags
    for cfg, cams in ((TrainConfig(iterations=3000, warmup_length=499), 64), (TrainConfig(), 64),
                      (TrainConfig(iterations=10000, warmup_length=500), 200), (TrainConfig(), 3000)):
        stop = cfg.iterations // 2
        for s in range(0, cfg.iterations + 1):
            assert refine_flags(s, cfg, cams, stop) == oracle.refine_flags(
                s, cfg.refine_every, cfg.warmup_length, cfg.reset_alpha_every, stop, cams,
                cfg.stop_screen_size_at), s


def test_core_surface_names_and_defaults():
    """msplat._core keeps the reference module's names and TrainingConfig defaults
    (python/bindings.cpp:494-527); a bad background is rejected as there."""
    import msplat
    from msplat import _core
    cfg = _core.TrainingConfig()
    want = dict(iterations=30000, sh_degree=3, sh_degree_interval=1000, ssim_weight=0.2, num_downscales=2,
                resolution_schedule=3000, refine_every=100, warmup_length=500, reset_alpha_every=30,
                densify_grad_thresh=0.0002, densify_size_thresh=0.01, stop_screen_size_at=4000, split_screen_size=0.05,
                keep_crs=False, downscale_factor=1.0, output="splat.ply", save_every=-1, depth_weight=0.0,
                depth_weight_final=-1.0, depth_huber_delta=0.05, depth_min_alpha=0.25, sky_alpha_weight=0.0,
                fill_weight=0.0, use_masks=False, learn_sky=False, sky_width=512, sky_height=128, sky_lr=0.01,
                exposure_compensation=False, exposure_lr=5e-3, exposure_reg=1e-2, max_scale_ratio=0.0)
    for k, v in want.items():
        assert getattr(cfg, k) == v, k
    assert cfg.bg_color == [0.6130, 0.0101, 0.3984] and not cfg.priors_enabled()
    with pytest.raises(ValueError):
        _core.TrainingConfig(bg_color=[1, 0])
    for name in ("TrainingConfig", "TrainingStats", "Dataset", "GaussianTrainer", "sync", "cleanup",
                 "_gpu_radix_sort", "_gpu_knn3_mean_dist", "_gpu_resize_area", "_gpu_undistort", "_gpu_image_metrics",
                 "_gpu_pack_rgba8"):
        assert callable(getattr(_core, name)), name
    assert msplat.TrainingConfig is _core.TrainingConfig


def test_undistort_roi_properties():
    """undistortROI (image_sizing.cpp:33-69): no distortion keeps the image less its last
    row and column (border pixel centres map to themselves); barrel distortion has
    nothing to crop beyond that; pincushion crops on every side."""
    from msplat._core import undistort_roi
    assert undistort_roi(120, 90, 100, 102, 60, 45, 0, 0, 0, 0, 0) == (0, 0, 119, 89)
    x, y, w, h = undistort_roi(120, 90, 100, 102, 61, 44, 0.25, 0.05, 0.001, -0.002, 0.01)
    assert x > 0 and y > 0 and x + w < 120 and y + h < 90


def test_prior_schedules_and_sky_frame():
    """Model::depthWeightAt (model.cpp:693-701): constant without a final weight, linear
    to zero with final 0, log-linear between two weights; computeSkyFrame (704-737): an
    orthonormal frame whose zenith is the mean camera up and whose azimuth 0 is the mean
    heading; initSkyTexture (739-812) on one camera looking along azimuth 0."""
    from msplat.camera import look_at, make_camera
    from msplat.trainer import PriorOptions, depth_weight_at, init_sky_texture, sky_frame
    o = PriorOptions(depth_weight=0.2)
    assert depth_weight_at(o, 0, 100) == np.float32(0.2) and depth_weight_at(o, 90, 100) == np.float32(0.2)
    o = PriorOptions(depth_weight=0.2, depth_weight_final=0.0)
    assert abs(depth_weight_at(o, 25, 100) - 0.15) < 1e-7 and depth_weight_at(o, 200, 100) == 0.0
    o = PriorOptions(depth_weight=0.2, depth_weight_final=0.05)
    assert abs(depth_weight_at(o, 50, 100) - 0.1) < 1e-6 and abs(depth_weight_at(o, 100, 100) - 0.05) < 1e-7
    assert depth_weight_at(PriorOptions(), 10, 100) == 0.0
    cams = [make_camera(look_at((0.0, 0.0, 0.0), (1.0, 0.02 * k, 0.0)), 64, 48, 50.0, 50.0, 32.0, 24.0)
            for k in range(-2, 3)]
    fr = sky_frame(cams).reshape(3, 3).astype(np.float64)
    np.testing.assert_allclose(fr @ fr.T, np.eye(3), atol=1e-6)
    np.testing.assert_allclose(fr[0], [0, 1, 0], atol=1e-3)   # zenith: the cameras' up
    np.testing.assert_allclose(fr[1], [1, 0, 0], atol=1e-3)   # azimuth 0: their heading
    cam = cams[2]
    img = np.zeros((48, 64, 3), np.float32)
    img[:24] = (0.2, 0.5, 0.9)                                 # sky above the horizon
    img[24:] = (0.3, 0.3, 0.3)
    aux = np.zeros((24, 32, 4), np.uint8)
    aux[:12, :, 1] = 255
    tex = init_sky_texture([cam], [img], [dict(aux=aux, has_sky=True)], sky_frame([cam]), 32, 16)
    assert tex.shape == (16, 32, 3)
    np.testing.assert_allclose(tex[:8].reshape(-1, 3), np.broadcast_to((0.2, 0.5, 0.9), (8 * 32, 3)), atol=1e-6)  # seen or row-filled
    np.testing.assert_allclose(tex[8:].reshape(-1, 3), np.broadcast_to((0.2, 0.5, 0.9), (8 * 32, 3)), atol=1e-6)
    blank = init_sky_texture([cam], [img], [None], sky_frame([cam]), 8, 4)
    np.testing.assert_allclose(blank, np.broadcast_to((0.6, 0.7, 0.85), (4, 8, 3)), atol=1e-7)
