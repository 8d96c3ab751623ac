"""HIP kernels (through the C ABI) against the oracle on the same seeded inputs.

Bars: integer outputs (radii, tile rectangles, per-tile sorted gaussian lists, tile
ranges) bit-exact; projection records bit-exact where no exp is involved (pixel centre,
depth) and within 2 ulp of each record's largest component where one is (conic,
colour, aabb: the kernels round exp once
from double, i.e. correctly, while glibc's expf — the reference's — is a 0.502-ulp
routine that differs on ~0.06% of inputs); the loss gradient bit-exact; composited
images, transmittance and gradients within 1e-4 relative (norm-wise per output array),
the composite differing from the CPU only through the hardware exp of the per-pixel loop.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import cam_from, golden

pytestmark = pytest.mark.gpu

KEYS = ("means", "scales", "quats", "dc", "rest", "opac")
BG = np.array([0.613, 0.0101, 0.3984], np.float32)


def _scene_cam(seed=1, n=3000, deg=3, W=97, H=71):
    from msplat.camera import look_at, make_camera
    from msplat.scenes import random_scene
    sc = random_scene(seed, n, degree=deg, log_scale=(-3.5, 0.5))
    cam = make_camera(look_at((2.5, 0.7, 2.9), (0.1, -0.2, 0)), W, H, 80.0, 83.0, W / 2 - 1.2, H / 2 + 0.9, dtu=deg)
    return sc, cam


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)


def _frame_arrays(ctx, n):
    f = ctx.frame()
    A = ctx.download(f.splat_a, (n, 4), np.float32)
    B = ctx.download(f.splat_b, (n, 4), np.float32)
    blue = ctx.download(f.splat_blue, (n,), np.float32)
    return dict(xys=A[:, :2], conics=np.c_[A[:, 2:4], B[:, 0]], opac=B[:, 1], colors=np.c_[B[:, 2:4], blue],
                depths=ctx.download(f.depths, (n,), np.float32), aabb=ctx.download(f.aabb, (n, 2), np.float32),
                radii=ctx.download(f.radii, (n,), np.int32), f=f)


@pytest.mark.parametrize("deg,dtu", [(0, 0), (1, 1), (2, 2), (3, 3), (3, 1), (4, 4), (4, 2)])
def test_projection_bit_exact(ctx, oracle, deg, dtu):
    """Projection and SH colour against the oracle, for every stored degree and fewer bands
    in use than stored (the colour reads the bands in use of each gaussian's stored ones:
    staged in LDS up to degree 3, read in place at degree 4)."""
    from msplat._abi import check, lib, make_camera_struct
    from msplat.backend import DeviceGaussians
    sc, cam = _scene_cam(seed=10 + deg, deg=deg)
    cam["dtu"] = dtu
    dg = DeviceGaussians(ctx, sc)
    g, c = dg.struct(), make_camera_struct(cam, deg)
    check(lib().msplat_hip_stage_project(ctx.ptr, C.byref(g), C.byref(c)))
    got = _frame_arrays(ctx, dg.n)
    ref = oracle.project(sc, cam)
    np.testing.assert_array_equal(got["radii"], ref["radii"])
    vis = ref["radii"] > 0
    assert vis.sum() > 100
    for k in ("xys", "depths", "colors"):  # no exp on their path
        np.testing.assert_array_equal(got[k][vis], ref[k][vis], err_msg=k)
    # conic = inverse 2D covariance: a relative change eps of the covariance (one ulp of an
    # exp) moves it by about kappa * eps, kappa the condition number
    a, b = got["conics"][vis].astype(np.float64), ref["conics"][vis].astype(np.float64)
    assert np.mean(np.any(a != b, axis=1)) < 0.01
    tr, det = b[:, 0] + b[:, 2], b[:, 0] * b[:, 2] - b[:, 1] ** 2
    disc = np.sqrt(np.maximum(tr * tr / 4 - det, 0))
    kappa = (tr / 2 + disc) / np.maximum(tr / 2 - disc, 1e-30)
    bound = 16 * kappa * 2.0 ** -24 * np.abs(b).max(axis=1)
    assert np.all(np.abs(a - b).max(axis=1) <= bound)
    # aabb = ceil(3 sqrt(cov.xx)): equal, or one apart where the ceil sits on a boundary
    a, b = got["aabb"][vis], ref["aabb"][vis]
    assert np.mean(a != b) < 0.01 and np.abs(a - b).max() <= 1


@pytest.mark.parametrize("ties", [False, True])
def test_binning_bit_exact_on_oracle_inputs(ctx, oracle, ties):
    """Keys, the stable sorts and the tile ranges, fed the oracle's projection. With ties,
    a third of the gaussians repeat another's mean (equal depth bits): the order of equal
    depths is the gaussian id order."""
    sc, cam = _scene_cam(seed=5, n=4000, deg=1, W=130, H=90)
    if ties:
        rng = np.random.default_rng(55)
        dst = rng.choice(4000, 1300, replace=False)
        sc["means"][dst] = sc["means"][rng.integers(0, 4000, 1300)]
    _check_binning(ctx, oracle, sc, cam)


def test_binning_bit_exact_with_huge_splats(ctx, oracle):
    """Key emission across its 2048-key workgroup slices: at 1920x1080 (8160 tiles) a few
    splats cover most of the image, so their keys span several slices and slices start
    inside them, among thousands of small splats whose records start anywhere in a
    thread's eight outputs."""
    from msplat.camera import look_at, make_camera
    from msplat.scenes import random_scene
    sc = random_scene(9, 4000, degree=1, log_scale=(-3.5, 0.5))
    sc["scales"][:6] = np.float32(0.2)
    cam = make_camera(look_at((2.5, 0.7, 2.9), (0.1, -0.2, 0)), 1920, 1080, 1500.0, 1500.0, dtu=1)
    M, counts = _check_binning(ctx, oracle, sc, cam)
    assert counts.max() > 3 * 2048 and (counts > 2048).sum() >= 3, counts.max()


def _check_binning(ctx, oracle, sc, cam):
    from msplat._abi import FRAME_KEYS64, check, lib, make_camera_struct
    from msplat.backend import DeviceGaussians
    dg = DeviceGaussians(ctx, sc)
    g, c = dg.struct(), make_camera_struct(cam, 1)
    check(lib().msplat_hip_stage_project(ctx.ptr, C.byref(g), C.byref(c)))
    ref = oracle.project(sc, cam)
    n = dg.n
    f = ctx.frame()
    sig = 1 / (1 + np.exp(-sc["opac"][:, 0].astype(np.float64)))
    ctx.upload(f.splat_a, np.c_[ref["xys"], ref["conics"][:, :2]].astype(np.float32))
    ctx.upload(f.splat_b, np.c_[ref["conics"][:, 2], sig, ref["colors"][:, :2]].astype(np.float32))
    ctx.upload(f.splat_blue, ref["colors"][:, 2].copy())
    ctx.upload(f.depths, ref["depths"])
    ctx.upload(f.aabb, ref["aabb"])
    ctx.upload(f.radii, ref["radii"])
    check(lib().msplat_hip_stage_bin(ctx.ptr, C.byref(c)))
    ctx.set_flags(FRAME_KEYS64)
    try:
        f = ctx.frame()
    finally:
        ctx.set_flags(0)
    bins = oracle.bin_sort(ref, sc["opac"], cam)
    M = bins["M"]
    assert f.num_isects == M and M > 1000
    gid = ctx.download(f.isect_gid, (M,), np.int32)
    keys = ctx.download(f.isect_keys, (M,), np.uint64)
    ts = ctx.download(f.tile_start, (cam["tiles_x"] * cam["tiles_y"] + 1,), np.int32)
    np.testing.assert_array_equal(gid, bins["isect_gid"])
    np.testing.assert_array_equal(ts[:-1], bins["tile_start"])
    np.testing.assert_array_equal(ts[1:], bins["tile_end"])
    np.testing.assert_array_equal((keys & np.uint64(0xFFFFFFFF)), bins["keys"] >> np.uint64(32))
    return M, np.bincount(bins["isect_gid"], minlength=n)  # intersections per gaussian


@pytest.mark.parametrize("deg,W,H", [(0, 97, 71), (3, 97, 71), (3, 256, 160)])
def test_render_matches_oracle(ctx, oracle, deg, W, H):
    from msplat.backend import DeviceGaussians, render_numpy
    sc, cam = _scene_cam(seed=20 + deg, n=6000, deg=deg, W=W, H=H)
    dg = DeviceGaussians(ctx, sc)
    img = render_numpy(ctx, dg, cam, BG)
    _, bins, fwd = oracle.render(sc, cam, BG)
    f = ctx.frame()
    assert f.num_isects == bins["M"]
    fT = ctx.download(f.final_T, (H, W), np.float32)
    fidx = ctx.download(f.final_idx, (H, W), np.int32)
    diff = np.abs(img - fwd["img"])
    print(f"\nrender deg{deg} {W}x{H}: M={bins['M']} rel={_rel(img, fwd['img']):.2e} max={diff.max():.2e} "
          f"idx-eq={np.mean(fidx == fwd['final_idx']):.5f}")
    assert _rel(img, fwd["img"]) < 1e-4
    assert _rel(fT, fwd["final_T"]) < 1e-4
    assert np.mean(diff > 1e-4) < 1e-3  # threshold flips of the per-pixel exp only
    assert np.mean(fidx == fwd["final_idx"]) > 0.999


# sizes around the loss kernels' 118-column x 8-row blocks and their 5-pixel halos
LOSS_SIZES = [(75, 53), (250, 37), (118, 16), (119, 9), (236, 8), (1, 1), (3, 300), (400, 5)]


@pytest.mark.parametrize("W,H", LOSS_SIZES)
def test_loss_bit_exact(ctx, oracle, W, H):
    from msplat._abi import DeviceArray, check, lib, make_camera_struct
    from msplat.backend import DeviceGaussians, render_numpy
    sc, cam = _scene_cam(seed=3, n=2000, deg=1, W=W, H=H)
    dg = DeviceGaussians(ctx, sc)
    render_numpy(ctx, dg, cam, BG)
    rng = np.random.default_rng(4 + W * H)
    img = rng.random((H, W, 3), dtype=np.float32)
    gt = np.clip(img + rng.normal(0, 0.1, img.shape), 0, 1).astype(np.float32)
    gt[::7, ::5] = img[::7, ::5]  # equal pixels: the L1 subgradient's zero branch
    # a target equal to the render on the left half: pixel losses of a few ulp there,
    # whose row sums the reference rounds (the kernels' in-order path)
    gt_half = gt.copy()
    gt_half[:, : (W + 1) // 2] = img[:, : (W + 1) // 2]
    f = ctx.frame()
    ctx.upload(f.image, img)
    c = make_camera_struct(cam, 1)
    for target in (gt, gt_half):
        dgt = DeviceArray.from_numpy(ctx, target)
        for w in (0.2, 0.0, 1.0):
            loss = C.c_float()
            inv_n = float(np.float32(1) / np.float32(img.size))
            check(lib().msplat_hip_stage_loss(ctx.ptr, C.byref(c), C.c_void_p(dgt.ptr), w, inv_n, C.byref(loss)))
            v = ctx.download(ctx.frame().v_image, img.shape, np.float32)
            tot, vref = oracle.loss(img, target, w)
            np.testing.assert_array_equal(v, vref, err_msg=f"{W}x{H} w {w}")
            dsum = ctx.download(ctx.frame().loss_sum, (1,), np.float64)[0]
            assert dsum == tot, (W, H, w, dsum, tot)  # the double sum, bit for bit
            assert np.float32(loss.value) == np.float32(tot / (H * W)), (W, H, w, loss.value, tot / (H * W))


def _oracle_grads(oracle, sc, cam, v_img):
    """Per-gaussian parameter gradients of <image, v_img> from the oracle."""
    p, b, fwd = oracle.render(sc, cam, BG)
    slots = oracle.raster_backward(b, fwd, cam, BG, v_img)
    n = sc["means"].shape[0]
    red = oracle.reduce_slots(n, b, slots)
    st = {k: np.ascontiguousarray(sc[k], np.float32).copy() for k in KEYS}
    vm, vs, vq = np.zeros((n, 3), np.float32), np.zeros((n, 3), np.float32), np.zeros((n, 4), np.float32)
    z = lambda a: np.zeros_like(a)
    oracle.lib.orc_f32_project_backward_adam(
        n, st["means"], st["scales"], 1.0, st["quats"], cam["viewmat"], cam["projmat"], cam["fx"], cam["fy"],
        cam["W"], cam["H"], sc["degree"], cam["dtu"], cam["cam_pos"], p["radii"], p["conics"], red["v_xy"],
        red["v_conic"], red["v_rgb"], red["v_depth"], vm, vs, vq, st["dc"].copy(), st["rest"].copy(), z(st["dc"]),
        z(st["dc"]), z(st["rest"]), z(st["rest"]), 0.0, 1.0, 0.0, 1.0, 0.9, 0.999, 1e-8)
    return red, vm, vs, vq


@pytest.mark.parametrize("ppt", [1, 2, 4])
def test_backward_gradients_match_oracle(ctx, oracle, ppt):
    from msplat._abi import DeviceArray, FRAME_KEEP_GRADS, TUNE_BWD_PPT, Grads, check, floats, lib, make_camera_struct
    from msplat.backend import DeviceGaussians, render
    sc, cam = _scene_cam(seed=8, n=3000, deg=3, W=97, H=71)
    dg = DeviceGaussians(ctx, sc)
    ctx.set_flags(FRAME_KEEP_GRADS)
    default_ppt = ctx.get_tuning(TUNE_BWD_PPT)
    ctx.set_tuning(TUNE_BWD_PPT, ppt)
    try:
        render(ctx, dg, cam, BG)
        rng = np.random.default_rng(9)
        v_img = (rng.normal(0, 1, (71, 97, 3)) * 1e-3).astype(np.float32)
        dv = DeviceArray.from_numpy(ctx, v_img)
        n = dg.n
        outs = {k: DeviceArray(ctx, s) for k, s in
                (("means", (n, 3)), ("scales", (n, 3)), ("quats", (n, 4)), ("dc", (n, 3)), ("rest", (n, 15, 3)),
                 ("opac", (n,)))}
        gr = Grads(outs["means"].ptr, outs["scales"].ptr, outs["quats"].ptr, outs["dc"].ptr, outs["rest"].ptr,
                   outs["opac"].ptr)
        g, c = dg.struct(), make_camera_struct(cam, 3)
        check(lib().msplat_hip_render_backward(ctx.ptr, C.byref(g), C.byref(c), floats(BG), C.c_void_p(dv.ptr),
                                               C.byref(gr)))
        f = ctx.frame()
        got = dict(v_xy=ctx.download(f.v_xy, (n, 2), np.float32), v_conic=ctx.download(f.v_conic, (n, 3), np.float32),
                   v_rgb=ctx.download(f.v_rgb, (n, 3), np.float32), v_opac=ctx.download(f.v_opacity, (n,), np.float32))
        red, vm, vs, vq = _oracle_grads(oracle, sc, cam, v_img)
        report = {}
        for k in ("v_xy", "v_conic", "v_rgb", "v_opac"):
            report[k] = _rel(got[k], red[k])
        report["means"] = _rel(outs["means"].numpy(), vm)
        report["scales"] = _rel(outs["scales"].numpy(), vs)
        report["quats"] = _rel(outs["quats"].numpy(), vq)
        report["opac"] = _rel(outs["opac"].numpy(), red["v_opac"])
        # SH bands: C0 v_rgb for dc, basis x v_rgb for the rest
        from msplat.camera import TILE  # noqa: F401
        d = sc["means"].astype(np.float64) - cam["cam_pos"]
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        x, y, z = d.T
        C1, C2, C3 = 0.4886025119029199, [1.0925484305920792, -1.0925484305920792, 0.31539156525252005,
                                          -1.0925484305920792, 0.5462742152960396], \
            [-0.5900435899266435, 2.890611442640554, -0.4570457994644658, 0.3731763325901154, -0.4570457994644658,
             1.445305721320277, -0.5900435899266435]
        basis = np.stack([-C1 * y, C1 * z, -C1 * x, C2[0] * x * y, C2[1] * y * z, C2[2] * (2 * z * z - x * x - y * y),
                          C2[3] * x * z, C2[4] * (x * x - y * y), C3[0] * y * (3 * x * x - y * y), C3[1] * x * y * z,
                          C3[2] * y * (4 * z * z - x * x - y * y), C3[3] * z * (2 * z * z - 3 * x * x - 3 * y * y),
                          C3[4] * x * (4 * z * z - x * x - y * y), C3[5] * z * (x * x - y * y),
                          C3[6] * x * (x * x - 3 * y * y)], 1)
        vis = (red["v_rgb"] != 0).any(1)[:, None, None]
        rest_ref = np.where(vis, basis[:, :, None] * red["v_rgb"][:, None, :], 0)
        report["dc"] = _rel(outs["dc"].numpy(), 0.28209479177387814 * red["v_rgb"])
        report["rest"] = _rel(outs["rest"].numpy(), rest_ref)
        print(f"\nbackward (ppt={ppt}) rel errors:", {k: f"{v:.2e}" for k, v in report.items()})
        for k, v in report.items():
            assert v < 1e-4, (k, v)
    finally:
        ctx.set_flags(0)
        ctx.set_tuning(TUNE_BWD_PPT, default_ppt)


def test_train_steps_match_golden(ctx):
    """Three GPU train steps against the reference's own three steps (golden vectors)."""
    from msplat._abi import DeviceArray
    from msplat.backend import DeviceGaussians, train_step
    d = golden("train_steps")
    sc = {k: d["init_" + k] for k in KEYS}
    sc["degree"] = 3
    dg = DeviceGaussians(ctx, sc)
    cams = [cam_from(d, "cam0_"), cam_from(d, "cam1_")]
    for t in range(3):
        gt = DeviceArray.from_numpy(ctx, d[f"gt{t}"])
        loss = train_step(ctx, dg, cams[t % 2], d["bg"], gt, d["steps"][t], d["bc2s"][t])
        assert abs(loss - d["losses"][t]) <= 1e-6 * abs(d["losses"][t]), (t, loss, d["losses"][t])
        radii = dg.radii.numpy()[:dg.n]
        if t == 0:  # identical inputs: identical radii
            np.testing.assert_array_equal(radii, d[f"radii{t}"])
        else:  # parameters now differ at rounding level; a ceil may land on the other side
            assert np.mean(radii == d[f"radii{t}"]) > 0.995
    st = dg.to_numpy()
    errs = {}
    for k in KEYS:
        for pre in ("", "m1_", "m2_"):
            ref = d["out_" + pre + k].reshape(st[pre + k].shape)
            errs[pre + k] = _rel(st[pre + k], ref)
    for k in ("vis", "gn", "ms"):
        errs[k] = _rel(st[k], d["out_" + k])
    print("\ntrain-step rel errors:", {k: f"{v:.1e}" for k, v in errs.items()})
    for k, v in errs.items():
        assert v < 1e-4, (k, v)


def test_train_steps_with_an_empty_frame_match_oracle(ctx, oracle):
    """A camera that sees no Gaussian (all behind it: no tile intersection), then one that
    sees them, then the empty one again. The empty frame renders the background, and
    Adam still runs on every Gaussian with zero gradients (the moments decay and the
    parameters move by the decayed first moment)."""
    from msplat._abi import DeviceArray
    from msplat.backend import DeviceGaussians, train_step
    from msplat.camera import look_at, make_camera
    sc, cam = _scene_cam(seed=31, n=2000, deg=3, W=97, H=71)
    away = make_camera(look_at((0.0, 0.0, 5.0), (0.0, 0.0, 10.0)), 97, 71, 80.0, 83.0, 47.3, 36.4, dtu=3)
    dg = DeviceGaussians(ctx, sc)
    st = {k: np.ascontiguousarray(sc[k], np.float32).copy() for k in KEYS}
    for k in KEYS:
        st["m1_" + k] = np.zeros_like(st[k])
        st["m2_" + k] = np.zeros_like(st[k])
    n = sc["means"].shape[0]
    st.update(degree=3, vis=np.zeros(n, np.float32), gn=np.zeros(n, np.float32), ms=np.zeros(n, np.float32))
    rng = np.random.default_rng(32)
    steps = np.array([1.6e-4, 5e-3, 1e-3, 2.5e-3, 1.25e-4, 5e-2], np.float32)
    for t, c in enumerate((away, cam, away)):
        gt_np = rng.uniform(0, 1, (71, 97, 3)).astype(np.float32)
        bc2 = np.full(6, np.float32(np.sqrt(1 - 0.999 ** (t + 1))), np.float32)
        loss = train_step(ctx, dg, c, BG, DeviceArray.from_numpy(ctx, gt_np), steps, bc2)
        ref_loss, ref_radii, ref_img = oracle.train_step(st, c, BG, gt_np, steps, bc2)
        f = ctx.frame()
        assert (f.num_isects == 0) == (c is away), (t, f.num_isects)
        assert abs(loss - ref_loss) <= 1e-6 * abs(ref_loss), (t, loss, ref_loss)
        radii = dg.radii.numpy()[:n]
        if c is away:
            assert not radii.any()
            np.testing.assert_array_equal(ref_img, np.broadcast_to(BG, ref_img.shape))
            np.testing.assert_array_equal(ctx.download(f.image, (71, 97, 3), np.float32), ref_img)
        else:
            assert np.mean(radii == ref_radii) > 0.995
    got = dg.to_numpy()
    errs = {}
    for k in KEYS:
        for pre in ("", "m1_", "m2_"):
            errs[pre + k] = _rel(got[pre + k], st[pre + k].reshape(got[pre + k].shape))
    for k in ("vis", "gn", "ms"):
        errs[k] = _rel(got[k], st[k])
    print("\nempty-frame train-step rel errors:", {k: f"{v:.1e}" for k, v in errs.items()})
    for k, v in errs.items():
        assert v < 1e-4, (k, v)


def test_train_step_at_street_scale_matches_oracle(ctx, oracle):
    """One train step at the headline size: 475,115 SH3 Gaussians of the street recipe
    (GPU kNN initialisation, which the oracle's matches bit for bit), a 1264x832 view on
    the street path, against the oracle on the same inputs (a few seconds of CPU). The
    initial log-scales are isotropic, which makes the rotation gradient zero up to
    rounding (noise on both sides), so they are perturbed per axis first. The SH rest
    coefficients get N(0, 0.05) values, as a trained scene has: from zero they would
    hold only Adam's first update after the step, step * g / (|g| + eps), whose relative
    error at a gaussian with a near-zero gradient is that gradient's own (a pair at the
    1/255 alpha cut-off, where the GPU's and the CPU's exp differ in the last ulp, moves
    it by percent), however small the gradient's error over the whole array."""
    from msplat._abi import DeviceArray
    from msplat.backend import DeviceGaussians, train_step
    from msplat.scenes import street_cameras, street_points
    n, W, H = 475_115, 1264, 832
    pts, rgb = street_points(0, n)
    dg = DeviceGaussians.from_points(ctx, pts, rgb, degree=3)
    scales = dg.params["scales"].numpy()
    scales[:n] += np.random.default_rng(37).normal(0, 0.3, (n, 3)).astype(np.float32)
    dg.params["scales"].upload(scales)
    rest = dg.params["rest"].numpy()
    rest[:n] = np.random.default_rng(38).normal(0, 0.05, (n, 45)).astype(np.float32)
    dg.params["rest"].upload(rest)
    st = dg.to_numpy()
    cam = dict(street_cameras(4, W, H, 900.0, 900.0)[1], dtu=3)
    gt_np = np.random.default_rng(36).uniform(0, 1, (H, W, 3)).astype(np.float32)
    steps = np.array([1.6e-4, 5e-3, 1e-3, 2.5e-3, 1.25e-4, 5e-2], np.float32)
    bc2 = np.full(6, np.float32(np.sqrt(1 - 0.999)), np.float32)
    loss = train_step(ctx, dg, cam, BG, DeviceArray.from_numpy(ctx, gt_np), steps, bc2)
    img = ctx.download(ctx.frame().image, (H, W, 3), np.float32)
    M = ctx.frame().num_isects
    ref_loss, ref_radii, ref_img = oracle.train_step(st, cam, BG, gt_np, steps, bc2)
    assert M > 1_000_000
    assert abs(loss - ref_loss) <= 1e-6 * abs(ref_loss), (loss, ref_loss)
    assert np.mean(dg.radii.numpy()[:n] == ref_radii) > 0.999
    assert _rel(img, ref_img) < 1e-4
    got = dg.to_numpy()
    errs = {}
    for k in KEYS:
        for pre in ("", "m1_", "m2_"):
            errs[pre + k] = _rel(got[pre + k], st[pre + k].reshape(got[pre + k].shape))
    for k in ("vis", "gn", "ms"):
        errs[k] = _rel(got[k], st[k])
    print(f"\nstreet-scale train step (M={M}) rel errors:", {k: f"{v:.1e}" for k, v in errs.items()})
    for k, v in errs.items():
        assert v < 1e-4, (k, v)


def test_train_step_without_gaussians_matches_oracle(ctx, oracle):
    """Zero Gaussians: the frame is the background and the loss is the oracle's."""
    from msplat._abi import DeviceArray
    from msplat.backend import DeviceGaussians, render_numpy, train_step
    from msplat.scenes import random_scene
    _, cam = _scene_cam(seed=1, n=1, deg=3)
    sc = random_scene(33, 0, degree=3)
    dg = DeviceGaussians(ctx, sc)
    np.testing.assert_array_equal(render_numpy(ctx, dg, cam, BG), np.broadcast_to(BG, (71, 97, 3)))
    gt_np = np.random.default_rng(34).uniform(0, 1, (71, 97, 3)).astype(np.float32)
    st = {k: np.ascontiguousarray(sc[k], np.float32).copy() for k in KEYS}
    for k in KEYS:
        st["m1_" + k] = np.zeros_like(st[k])
        st["m2_" + k] = np.zeros_like(st[k])
    st.update(degree=3, vis=np.zeros(0, np.float32), gn=np.zeros(0, np.float32), ms=np.zeros(0, np.float32))
    steps, bc2 = np.full(6, 1e-3, np.float32), np.full(6, 0.5, np.float32)
    loss = train_step(ctx, dg, cam, BG, DeviceArray.from_numpy(ctx, gt_np), steps, bc2)
    ref_loss, _, _ = oracle.train_step(st, cam, BG, gt_np, steps, bc2)
    assert ctx.frame().num_isects == 0
    assert loss == np.float32(ref_loss), (loss, ref_loss)


@pytest.mark.parametrize("n,bits,hi", [(1, 32, 10), (5, 32, 4), (1023, 32, 2 ** 30), (1024, 32, 2 ** 30),
                                       (1025, 32, 7), (70_000, 32, 2 ** 30), (5_000, 64, 2 ** 63), (2_049, 64, 3),
                                       (300_000, 45, 2 ** 45), (50_000, 30, 2 ** 34), (3_000, 45, 2 ** 63),
                                       (4_096, 5, 2 ** 12)])
def test_radix_sort_matches_numpy(ctx, n, bits, hi):
    """The reference's own radix-sort cases (tests/test_msplat.py:429-441), a large one, and
    keys with bits set above key_bits, which the reference ignores when it orders them
    (core/cpu/msplat_cpu.cpp:1484-1486: the comparison masks both keys)."""
    from msplat._abi import DeviceArray, check, lib
    rng = np.random.default_rng(n + bits)
    keys = rng.integers(0, hi, n, dtype=np.uint64)
    dk = DeviceArray.from_numpy(ctx, keys)
    dv = DeviceArray.from_numpy(ctx, np.arange(n, dtype=np.int32))
    check(lib().msplat_hip_radix_sort(ctx.ptr, C.c_void_p(dk.ptr), C.c_void_p(dv.ptr), n, bits))
    order = np.argsort(keys & np.uint64((1 << bits) - 1 if bits < 64 else 0xFFFFFFFFFFFFFFFF), kind="stable")
    np.testing.assert_array_equal(dk.numpy(), keys[order])
    np.testing.assert_array_equal(dv.numpy(), order)


def test_written_slot_count_matches_oracle(ctx, oracle):
    """The raster backward writes (and counts) exactly the slots whose gradient is not
    all zero; the others are only flagged. With a cotangent that is nonzero everywhere,
    a slot is nonzero exactly when its pair composites some pixel, so the count equals
    the number of nonzero rows of the oracle's slots."""
    from msplat._abi import DeviceArray
    from msplat.backend import DeviceGaussians, render, render_backward
    sc, cam = _scene_cam(seed=21, n=4000, deg=3, W=97, H=71)
    dg = DeviceGaussians(ctx, sc)
    rng = np.random.default_rng(22)
    v_img = (rng.normal(0, 1, (71, 97, 3)) * 1e-3).astype(np.float32)
    render(ctx, dg, cam, BG)
    f = ctx.frame()
    before = int(ctx.download(f.slots_written, (1,), np.uint64)[0])
    render_backward(ctx, dg, cam, BG, DeviceArray.from_numpy(ctx, v_img))
    written = int(ctx.download(ctx.frame().slots_written, (1,), np.uint64)[0]) - before
    _, b, fwd = oracle.render(sc, cam, BG)
    slots = oracle.raster_backward(b, fwd, cam, BG, v_img)
    nonzero = int(np.any(slots != 0, axis=1).sum())
    assert b["M"] == ctx.frame().num_isects
    assert 0 < nonzero < b["M"]
    assert written == nonzero, (written, nonzero, b["M"])
