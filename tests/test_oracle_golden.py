"""The oracle (oracle/msplat_oracle.cpp) against the golden vectors written by the
compiled reference CPU backend (tests/golden/make_golden.py), bit for bit, and against
closed forms and brute force where the mathematics fixes the answer."""
import numpy as np
import pytest

from conftest import cam_from, golden

KEYS = ("means", "scales", "quats", "dc", "rest", "opac")


@pytest.mark.parametrize("deg", [0, 1, 3])
def test_render_matches_reference(oracle, deg):
    d = golden(f"render_deg{deg}")
    scene = {k: d[k] for k in KEYS}
    scene["degree"] = deg
    for i in range(2):
        cam = cam_from(d, f"cam{i}_")
        _, bins, fwd = oracle.render(scene, cam, d["bg"])
        assert bins["M"] > 0
        np.testing.assert_array_equal(fwd["img"], d[f"img{i}"])
        np.testing.assert_array_equal(fwd["final_T"], d[f"final_T{i}"])


def _train_state(d):
    st = {k: d["init_" + k].copy() for k in KEYS}
    for k in KEYS:
        st["m1_" + k] = np.zeros_like(st[k])
        st["m2_" + k] = np.zeros_like(st[k])
    n = st["means"].shape[0]
    st.update(degree=3, vis=np.zeros(n, np.float32), gn=np.zeros(n, np.float32), ms=np.zeros(n, np.float32))
    return st


def test_train_steps_match_reference(oracle):
    d = golden("train_steps")
    st = _train_state(d)
    cams = [cam_from(d, "cam0_"), cam_from(d, "cam1_")]
    for t in range(3):
        loss, radii, _ = oracle.train_step(st, cams[t % 2], d["bg"], d[f"gt{t}"], d["steps"][t], d["bc2s"][t])
        assert np.float32(loss) == d["losses"][t]
        np.testing.assert_array_equal(radii, d[f"radii{t}"])
    for k in KEYS:
        for pre in ("", "m1_", "m2_"):
            np.testing.assert_array_equal(st[pre + k], d["out_" + pre + k], err_msg=pre + k)
    for k in ("vis", "gn", "ms"):
        np.testing.assert_array_equal(st[k], d["out_" + k], err_msg=k)


@pytest.mark.parametrize("huge", [0, 1])
def test_densify_matches_reference(oracle, huge):
    d = golden("densify")
    N, fr = int(d["N"]), int(d["fr"])
    cap = 3 * N
    bufs = []
    for i in range(18):
        b = np.zeros((cap, d[f"in{i}"].shape[1]), np.float32)
        b[:N] = d[f"in{i}"]
        bufs.append(b)
    n, split, dup, keep = oracle.densify(bufs, N, cap, d["gn"], d["vc"], d["ms"], 40.0, fr, int(d["seed"]),
                                         cull_alpha=0.1, check_huge=huge)
    assert n == int(d[f"count_huge{huge}"])
    assert split.sum() > 0 and dup.sum() > 0 and keep.sum() == n
    for i in range(18):
        np.testing.assert_array_equal(bufs[i][:n], d[f"out{huge}_{i}"], err_msg=f"buffer {i}")


def test_opacity_reset_matches_reference(oracle):
    d = golden("opacity_reset")
    o, m1, m2 = d["o"].copy(), d["m1"].copy(), d["m2"].copy()
    oracle.opacity_reset(o, m1, m2, d["reset"])
    np.testing.assert_array_equal(o, d["out_o"])
    np.testing.assert_array_equal(m1, d["out_m1"])
    np.testing.assert_array_equal(m2, d["out_m2"])


def test_opacity_reset_closed_form(oracle):
    """The reset caps every opacity at 0.2 (logit -1.3862944 as a float32) and restarts
    the opacity group's Adam moments: o' = min(o, logit 0.2), m1 = m2 = 0. Values at or
    below the cap, and -inf, are left bit-identical; +inf and NaN-free inputs above it
    become the cap exactly."""
    g = np.random.default_rng(5)
    reset = np.float32(np.log(0.2 / 0.8))
    assert reset == np.float32(-1.3862943611198906)
    o = np.concatenate([g.normal(0, 4, 4000), [reset, -np.inf, np.inf, 1e30, -1e30]]).astype(np.float32)
    m1, m2 = g.normal(0, 1, o.size).astype(np.float32), g.random(o.size).astype(np.float32)
    before = o.copy()
    oracle.opacity_reset(o, m1, m2, reset)
    np.testing.assert_array_equal(o, np.minimum(before, reset))
    assert not np.any(m1) and not np.any(m2)
    # in probability space: sigmoid(o') = min(sigmoid(o), 0.2) to float rounding
    p = 1 / (1 + np.exp(-o.astype(np.float64)))
    np.testing.assert_allclose(p, np.minimum(1 / (1 + np.exp(-before.astype(np.float64))), 0.2), rtol=1e-6)


def test_knn_and_init_match_reference(oracle):
    d = golden("knn_init")
    np.testing.assert_array_equal(oracle.knn3(d["points"]), d["knn"])
    s, q, dc, o = oracle.init_gaussians(d["points"], d["rgb"], seed=42)
    for got, key in ((s, "scales"), (q, "quats"), (dc, "dc"), (o, "opac")):
        np.testing.assert_array_equal(got, d[key], err_msg=key)


def test_radix_sort_matches_reference_and_numpy(oracle):
    d = golden("radix_sort")
    for i in range(5):
        keys, bits = d[f"keys{i}"], int(d[f"bits{i}"])
        k, v = oracle.radix_sort(keys, np.arange(keys.size, dtype=np.int32), bits)
        np.testing.assert_array_equal(k, d[f"sorted{i}"])
        np.testing.assert_array_equal(v, d[f"order{i}"])
        mask = np.uint64((1 << bits) - 1) if bits < 64 else np.uint64(0xFFFFFFFFFFFFFFFF)
        np.testing.assert_array_equal(v, np.argsort(keys & mask, kind="stable"))


def test_image_ops_match_reference(oracle):
    d = golden("image_ops")
    np.testing.assert_array_equal(oracle.image_metrics(d["a"], d["b"]), d["metrics"])
    np.testing.assert_array_equal(oracle.pack_rgba8(d["img"]), d["rgba"])


# ── closed forms and brute force ─────────────────────────────────────────────

def _pcg(v):
    m = 0xFFFFFFFF
    st = (v * 747796405 + 2891336453) & m
    w = (((st >> ((st >> 28) + 4)) ^ st) * 277803737) & m
    return ((w >> 22) ^ w) & m


def test_pcg_known_answers(oracle):
    for v in (0, 1, 42, 123456789, 0xFFFFFFFF):
        assert oracle.pcg_hash(v) == _pcg(v)
    u = oracle.rand_uniform(7, 3)
    h = _pcg(7 ^ _pcg(3))
    assert u == np.float32((np.float32(h >> 8) + np.float32(0.5)) * np.float32(1 / 16777216))
    n = oracle.rand_normal3(7, 5)
    us = [np.float64(oracle.rand_uniform(7, 20 + i)) for i in range(4)]
    ref = [np.sqrt(-2 * np.log(us[0])) * np.cos(2 * np.pi * us[1]), np.sqrt(-2 * np.log(us[0])) * np.sin(2 * np.pi * us[1]),
           np.sqrt(-2 * np.log(us[2])) * np.cos(2 * np.pi * us[3])]
    np.testing.assert_allclose(n, ref, rtol=1e-5, atol=1e-6)


def _axis_camera(W=64, H=48, f=50.0):
    from msplat.camera import look_at, make_camera
    return make_camera(look_at((0, 0, 5), (0, 0, 0)), W, H, f, f, W / 2, H / 2)


def test_projection_closed_form(oracle):
    """An isotropic gaussian on the optical axis: 2D covariance f^2 s^2 / z^2 + 0.3,
    conic its inverse, radius ceil(3 sqrt(b + sqrt(max(0.1, b^2 - det)))) with b the
    mean eigenvalue (b^2 = det here, so the 0.1 floor applies), colour C0 * dc."""
    cam = _axis_camera()
    s = 0.08
    scene = dict(means=np.zeros((1, 3)), scales=np.full((1, 3), np.log(s)), quats=np.array([[0.3, 0.5, -0.2, 0.7]]),
                 dc=np.array([[0.4, -0.2, 0.1]]), rest=np.zeros((1, 1, 3)), opac=np.zeros((1, 1)), degree=0)
    p = oracle.project(scene, cam, dtype=np.float64)
    var = (50.0 * s / 5.0) ** 2 + 0.3
    np.testing.assert_allclose(p["conics"][0], [1 / var, 0, 1 / var], rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(p["depths"][0], 5.0, rtol=1e-12)
    assert p["radii"][0] == int(np.ceil(3 * np.sqrt(var + np.sqrt(0.1))))
    np.testing.assert_allclose(p["xys"][0], [cam["cx"] - 0.5, cam["cy"] - 0.5], atol=1e-9)
    np.testing.assert_allclose(p["colors"][0], 0.28209479177387814 * scene["dc"][0], rtol=1e-12)


def test_projection_rotation_invariant(oracle):
    """Rotating an anisotropic gaussian keeps its 2D covariance's determinant when the
    rotation is about the optical axis by 90 degrees (the ellipse rotates)."""
    cam = _axis_camera()
    q0 = np.array([1.0, 0, 0, 0])
    qz = np.array([np.cos(np.pi / 4), 0, 0, np.sin(np.pi / 4)])  # 90 degrees about z
    base = dict(means=np.zeros((2, 3)), scales=np.log([[0.05, 0.12, 0.07], [0.05, 0.12, 0.07]]),
                quats=np.stack([q0, qz]), dc=np.zeros((2, 3)), rest=np.zeros((2, 1, 3)), opac=np.zeros((2, 1)),
                degree=0)
    p = oracle.project(base, cam, dtype=np.float64)
    c0, c1 = p["conics"]
    np.testing.assert_allclose([c0[0], c0[2]], [c1[2], c1[0]], rtol=1e-9)


def test_compositing_closed_form(oracle):
    """Two gaussians centred on one pixel: C = a1 c1 + (1 - a1) a2 c2 + (1 - a1)(1 - a2) bg
    with a = min(0.999, sigmoid(o)) and c = max(raw + 0.5, 0); the front one first."""
    cam = _axis_camera(W=32, H=32, f=40.0)
    means = np.array([[0, 0, 0], [0, 0, -1.0]])  # the second one is further away
    o = np.array([[0.3], [-0.4]])
    dc = np.array([[0.2, -0.1, 0.3], [-0.3, 0.4, 0.0]])
    scene = dict(means=means, scales=np.full((2, 3), np.log(0.2)), quats=np.tile([1.0, 0, 0, 0], (2, 1)), dc=dc,
                 rest=np.zeros((2, 1, 3)), opac=o, degree=0)
    bg = np.array([0.1, 0.5, 0.9])
    _, _, fwd = oracle.render(scene, cam, bg, dtype=np.float64)
    a = np.minimum(0.999, 1 / (1 + np.exp(-o[:, 0])))
    c = np.maximum(0.28209479177387814 * dc + 0.5, 0)
    # the pixel at the projected centre (cx - 0.5, cy - 0.5) = (15.5, 15.5) is between pixels;
    # evaluate the exact gaussian weights at pixel (15, 15) instead
    p = oracle.project(scene, cam, dtype=np.float64)
    w = []
    for i in range(2):
        dx, dy = p["xys"][i] - 15.0
        cx, cy, cz = p["conics"][i]
        w.append(min(0.999, a[i] * np.exp(-(0.5 * (cx * dx * dx + cz * dy * dy) + cy * dx * dy))))
    ref = w[0] * c[0] + (1 - w[0]) * w[1] * c[1] + (1 - w[0]) * (1 - w[1]) * bg
    np.testing.assert_allclose(fwd["img"][15, 15], ref, rtol=1e-12)
    np.testing.assert_allclose(fwd["final_T"][15, 15], (1 - w[0]) * (1 - w[1]), rtol=1e-12)


def test_loss_closed_forms(oracle):
    rng = np.random.default_rng(0)
    img = rng.random((20, 27, 3), dtype=np.float32)
    gt = rng.random((20, 27, 3), dtype=np.float32)
    # pure L1: loss = mean |r - g|, gradient sign(r - g) / n
    tot, v = oracle.loss(img, gt, ssim_weight=0.0)
    n = img.size
    np.testing.assert_allclose(tot / (20 * 27), np.abs(img - gt).mean(), rtol=1e-5)
    np.testing.assert_array_equal(v, (np.float32(1.0) / np.float32(n)) * np.sign(img - gt).astype(np.float32))
    # SSIM of an image with itself is 1 everywhere: zero loss, zero gradient
    tot, v = oracle.loss(img, img, ssim_weight=1.0)
    assert abs(tot) < 1e-4
    assert np.abs(v).max() < 1e-6


def test_loss_gradient_finite_difference(oracle):
    """dL/d(render) of the L1 + SSIM loss against central differences of the float loss."""
    rng = np.random.default_rng(1)
    H, W = 14, 19
    gt = rng.random((H, W, 3), dtype=np.float32)
    img = np.clip(gt + rng.normal(0, 0.2, gt.shape), 0, 1).astype(np.float32)
    tot, v = oracle.loss(img, gt, ssim_weight=0.8)
    h = 1e-2
    for (y, x, c) in [(0, 0, 0), (7, 9, 1), (13, 18, 2), (3, 15, 0)]:
        p, m = img.copy(), img.copy()
        # stay away from the L1 kink
        if abs(img[y, x, c] - gt[y, x, c]) < 2 * h:
            continue
        p[y, x, c] += h
        m[y, x, c] -= h
        fd = (oracle.loss(p, gt, 0.8)[0] - oracle.loss(m, gt, 0.8)[0]) / (2 * h) / (H * W)
        np.testing.assert_allclose(v[y, x, c] * 3, fd * 3, rtol=2e-2, atol=2e-5)


def test_knn_bruteforce(oracle):
    rng = np.random.default_rng(7)
    pts = np.vstack([rng.uniform(-1, 1, (400, 3)), np.repeat(rng.uniform(-1, 1, (5, 3)), 3, axis=0)]).astype(np.float32)
    d2 = ((pts[:, None, :].astype(np.float64) - pts[None, :, :]) ** 2).sum(-1)
    np.fill_diagonal(d2, np.inf)
    ref = np.sqrt(np.sort(d2, axis=1)[:, :3]).mean(axis=1)
    np.testing.assert_allclose(oracle.knn3(pts), ref, rtol=1e-5, atol=1e-6)


def _knn_tiny_expected(pts):
    """Mean of the up-to-3 nearest distances; 1.0 for a point with no neighbour
    (core/cpu/msplat_cpu.cpp:1572-1576)."""
    d2 = ((pts[:, None, :].astype(np.float64) - pts[None, :, :]) ** 2).sum(-1)
    np.fill_diagonal(d2, np.inf)
    k = min(3, pts.shape[0] - 1)
    return np.sqrt(np.sort(d2, axis=1)[:, :k]).mean(axis=1) if k else np.ones(pts.shape[0])


@pytest.mark.parametrize("n", [1, 2, 3, 4])
def test_knn_fewer_than_three_neighbours(oracle, n):
    pts = np.random.default_rng(n).uniform(-1, 1, (n, 3)).astype(np.float32)
    np.testing.assert_allclose(oracle.knn3(pts), _knn_tiny_expected(pts), rtol=1e-6, atol=1e-7)


def test_render_double_agrees_with_float(oracle):
    """The FP64 renderer agrees with the FP32 one to float rounding on a smooth scene."""
    from msplat.scenes import random_scene
    from msplat.camera import look_at, make_camera
    sc = random_scene(3, 200, degree=1, log_scale=(-2.5, 0.3))
    cam = make_camera(look_at((2.5, 0.5, 2.5), (0, 0, 0)), 48, 40, 40.0, 40.0, dtu=1)
    bg = np.array([0.2, 0.3, 0.4], np.float32)
    _, _, f32 = oracle.render(sc, cam, bg)
    _, _, f64 = oracle.render(sc, cam, bg, dtype=np.float64)
    err = np.abs(f32["img"] - f64["img"])
    assert np.quantile(err, 0.999) < 1e-5


def test_image_pipeline_matches_reference(oracle):
    """msplat_resize_area and msplat_undistort: the oracle against the compiled
    reference's outputs, bit for bit."""
    from workloads import RESIZE_CASES, UNDISTORT_CASES
    d = golden("image_pipeline")
    for i, (_, sw, sh, dw, dh) in enumerate(RESIZE_CASES):
        np.testing.assert_array_equal(oracle.resize_area(d[f"resize_src{i}"], dw, dh), d[f"resize_out{i}"])
    for i, (intr, dist, rx, ry, dw, dh) in enumerate(UNDISTORT_CASES):
        got = oracle.undistort(d["undistort_src"], np.array(intr, np.float32), np.array(dist, np.float32), rx, ry,
                               dw, dh)
        np.testing.assert_array_equal(got, d[f"undistort_out{i}"])


def test_image_pipeline_closed_forms(oracle):
    """Resizing by whole factors is the block mean; equal sizes convert (u8 / 255); a
    constant image stays constant at any size; undistortion without distortion at roi 0
    reproduces the image (bilinear samples at the source pixel centres)."""
    g = np.random.default_rng(3)
    img = g.random((24, 36, 3), dtype=np.float32)
    block = img.reshape(8, 3, 12, 3, 3).astype(np.float64).mean(axis=(1, 3))
    np.testing.assert_allclose(oracle.resize_area(img, 12, 8), block, rtol=2e-6, atol=2e-7)
    u8 = g.integers(0, 256, (9, 7, 4), dtype=np.uint8)
    np.testing.assert_array_equal(oracle.resize_area(u8, 7, 9), u8[..., :3].astype(np.float32) * np.float32(1 / 255))
    const = np.full((13, 17, 3), 0.375, np.float32)
    for dw, dh in ((5, 4), (17, 13), (40, 29)):
        np.testing.assert_allclose(oracle.resize_area(const, dw, dh), 0.375, rtol=1e-6)
    intr = np.array([50.0, 50.0, 18.0, 12.0], np.float32)
    out = oracle.undistort(img, intr, np.zeros(5, np.float32), 0, 0, 36, 24)
    np.testing.assert_allclose(out, img, rtol=0, atol=1e-6)
