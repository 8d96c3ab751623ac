"""The oracle's FP32 backward chain (raster backward, slot reduction, projection/SH
backward) against central finite differences (h = 1e-3) of its own FP64 renderer,
on the smooth 64-Gaussian scene. This pins the oracle's gradients by the mathematics,
independently of the reference's golden vectors.

Gradients are read from the first Adam moment of one train step with beta1 = 0 and a
zero step size (m1 = g exactly; as the reference's tests/gradcheck.cpp reads them through
Adam). The loss is the training loss w (1 - SSIM) + (1 - w) L1."""
import numpy as np
import pytest

import gradcheck_scene as gs


@pytest.mark.parametrize("view_dependent", [False, True])
def test_oracle_backward_matches_fp64_finite_differences(oracle, view_dependent):
    sc, cam = gs.scene(view_dependent), gs.camera()
    st = {k: sc[k].copy() for k in gs.KEYS}
    st["degree"] = sc["degree"]
    for k in gs.KEYS:
        st["m1_" + k] = np.zeros_like(sc[k])
        st["m2_" + k] = np.zeros_like(sc[k])
    n = sc["means"].shape[0]
    st["vis"], st["gn"], st["ms"] = (np.zeros(n, np.float32) for _ in range(3))
    gt = gs.target(oracle, sc, cam).astype(np.float32)
    loss, radii, img = oracle.train_step(st, cam, gs.BG.astype(np.float32), gt, np.zeros(6, np.float32),
                                         np.ones(6, np.float32), ssim_weight=gs.SSIM_WEIGHT, b1=0.0)
    for k in gs.KEYS:  # zero step: the parameters are untouched
        np.testing.assert_array_equal(st[k], sc[k])
    assert abs(loss - gs.train_loss(img, gt)) < 1e-6
    grads = {k: st["m1_" + k] for k in gs.KEYS}
    fds = gs.finite_differences(oracle, sc, cam, "train")
    assert gs.compare(grads, fds, f"oracle view_dependent={view_dependent}") >= 64 * 20


def test_gradcheck_scene_is_smooth(oracle):
    """The margins the finite differences rely on: alpha above 1/255 and T above 1e-4
    on every pixel, every Gaussian covering every tile."""
    for vd in (False, True):
        sc, cam = gs.scene(vd), gs.camera()
        p, _, r = oracle.render(sc, cam, gs.BG, dtype=np.float64)
        yy, xx = np.mgrid[0:32, 0:32] + 0.5
        dx = xx[None] - p["xys"][:, 0, None, None]
        dy = yy[None] - p["xys"][:, 1, None, None]
        c = p["conics"]
        sigma = 0.5 * (c[:, 0, None, None] * dx * dx + c[:, 2, None, None] * dy * dy) + c[:, 1, None, None] * dx * dy
        alpha = 1 / (1 + np.exp(-sc["opac"][:, 0, None, None].astype(np.float64))) * np.exp(-sigma)
        assert alpha.min() > 4 / 255 and alpha.max() < 0.5
        assert r["final_T"].min() > 1e-3
        tile = (yy.astype(int) // 16) * 2 + xx.astype(int) // 16
        assert (r["final_idx"] == 64 * tile + 63).all()  # every pixel composites all 64 of its tile
        assert p["radii"].min() > 45  # every rect covers the 2x2 tiles


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_oracle_analytic_gradients_match_fp64_finite_differences(oracle, dtype):
    """Oracle.gradients (the analytic backward of <image, R>, in double for the FP64
    renderer's per-output errors and in float) against central differences of the FP64
    renderer, view-independent colours (means included)."""
    sc, cam = gs.scene(False), gs.camera()
    R = gs.linear_weights()
    g = oracle.gradients(sc, cam, gs.BG, R, dtype=dtype)
    assert g["img"].dtype == dtype
    fds = gs.finite_differences(oracle, sc, cam, "linear")
    assert gs.compare({k: g[k] for k in gs.KEYS}, fds, f"analytic {np.dtype(dtype).name}") >= 64 * 20
