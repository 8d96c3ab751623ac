"""Prior-guided training on the GPU (msplat_hip_train_step_prior / _render_prior /
_render_aux_outputs; PriorStep, core/metal/bindings.h:41-77) against the oracle's
prior path, which reproduces the compiled reference bit for bit
(tests/test_oracle_prior.py): every feature combination of tests/workloads.py, two
steps each, images, expected depth, transmittance, parameters, moments, statistics,
sky texture and moments, exposure rows and moments and the loss terms within the 1e-4
relative bar; the aux render; a larger view."""
import numpy as np
import pytest

from workloads import BG, PRIOR_VARIANTS, prior_workload

pytestmark = pytest.mark.gpu

KEYS = ("means", "scales", "quats", "dc", "rest", "opac")
STEPS = np.array([1.6e-4, 5e-3, 1e-3, 2.5e-3, 1.25e-4, 5e-2], np.float32)
BC2 = np.full(6, np.float32(np.sqrt(np.float32(1 - 0.999))), np.float32)
PRIOR_OUT = ("sky_tex", "sky_grad", "sky_m1", "sky_m2", "loss_terms", "expo_params", "expo_m1", "expo_m2",
             "expo_grad")


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def _state(sc):
    st = {k: np.ascontiguousarray(sc[k], np.float32).copy() for k in KEYS}
    for k in KEYS:
        st["m1_" + k] = np.zeros_like(st[k])
        st["m2_" + k] = np.zeros_like(st[k])
    n = st["means"].shape[0]
    st.update(degree=sc["degree"], vis=np.zeros(n, np.float32), gn=np.zeros(n, np.float32),
              ms=np.zeros(n, np.float32))
    return st


def _compare(ctx, oracle, variant, seed=3, n=600, W=64, H=48, steps=2, sky=(16, 32)):
    from msplat._abi import DeviceArray
    from msplat.backend import DeviceGaussians
    from msplat.prior import DevicePrior, aux_outputs, train_step_prior
    sc, cam, gt, prior = prior_workload(seed=seed, n=n, W=W, H=H, features=PRIOR_VARIANTS[variant], sky=sky)
    dg = DeviceGaussians(ctx, sc)
    dp = DevicePrior(ctx, prior)
    gt_d = DeviceArray.from_numpy(ctx, gt)
    st = _state(sc)
    po = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in prior.items()}
    errs = {}
    for t in range(steps):
        loss = train_step_prior(ctx, dg, cam, BG, gt_d, STEPS, BC2, dp)
        ref_loss, ref_radii, ref_img, ref_depth, ref_T = oracle.train_step_prior(st, cam, BG, gt, STEPS, BC2, po)
        errs[f"loss{t}"] = abs(loss - ref_loss) / abs(ref_loss)
        img = ctx.download(ctx.frame().image, (H, W, 3), np.float32)
        errs[f"img{t}"] = _rel(img, ref_img)
        assert np.mean(dg.radii.numpy()[:dg.n] == ref_radii) > 0.995
        if prior.get("aux"):
            depth, fT = aux_outputs(ctx, H, W)
            errs[f"depth{t}"] = _rel(depth, ref_depth)
            errs[f"final_T{t}"] = _rel(fT, ref_T)
    got = dg.to_numpy()
    for k in KEYS:
        for pre in ("", "m1_", "m2_"):
            errs[pre + k] = _rel(got[pre + k], st[pre + k].reshape(got[pre + k].shape))
    for k in ("vis", "gn", "ms"):
        errs[k] = _rel(got[k], st[k])
    for k in PRIOR_OUT:
        if k in dp.arrays:
            errs["prior_" + k] = _rel(dp.numpy(k), po[k])
    return errs


@pytest.mark.parametrize("variant", sorted(PRIOR_VARIANTS))
def test_prior_steps_match_oracle(ctx, oracle, variant):
    errs = _compare(ctx, oracle, variant)
    worst = max(errs, key=errs.get)
    print(f"\n{variant}: worst {worst} {errs[worst]:.2e};", {k: f"{v:.1e}" for k, v in errs.items() if v > 1e-6})
    for k, v in errs.items():
        assert v < 1e-4, (k, v)


def test_prior_steps_at_a_larger_view_match_oracle(ctx, oracle):
    """Every feature on 20,000 gaussians in a 320x240 view (300 tiles)."""
    errs = _compare(ctx, oracle, "all", seed=11, n=20000, W=320, H=240)
    worst = max(errs, key=errs.get)
    print(f"\n320x240: worst {worst} {errs[worst]:.2e}")
    for k, v in errs.items():
        assert v < 1e-4, (k, v)


def test_prior_steps_with_a_full_sky_texture_match_oracle(ctx, oracle):
    """Every feature on 20,000 gaussians in a 640x416 view with the default 512x128 sky: a
    texel spans about 8 pixels across, so each wave of the sky gradient sums many pixels
    per texel before its atomics (the texture's gradient, moments and values are checked)."""
    errs = _compare(ctx, oracle, "all", seed=12, n=20000, W=640, H=416, sky=(128, 512))
    worst = max(errs, key=errs.get)
    print(f"\n640x416, 512x128 sky: worst {worst} {errs[worst]:.2e}")
    for k, v in errs.items():
        assert v < 1e-4, (k, v)


@pytest.mark.parametrize("variant", ["all", "aux_sky", "aux_defaults"])
def test_prior_render_matches_oracle(ctx, oracle, variant):
    from msplat.backend import DeviceGaussians
    from msplat.prior import DevicePrior, aux_outputs, render_prior
    sc, cam, gt, prior = prior_workload(seed=5, features=PRIOR_VARIANTS[variant])
    dg = DeviceGaussians(ctx, sc)
    img = render_prior(ctx, dg, cam, BG, DevicePrior(ctx, prior))
    depth, fT = aux_outputs(ctx, cam["H"], cam["W"])
    ri, rd, rf, _ = oracle.render_prior(sc, cam, BG, prior)
    assert _rel(img, ri) < 1e-5 and _rel(depth, rd) < 1e-5 and _rel(fT, rf) < 1e-5
    # the sky shows where the gaussians leave the view transparent
    if prior.get("sky"):
        clear = fT > 0.99
        assert clear.sum() > 50
        assert np.abs(img[clear] - np.asarray(BG, np.float32)).max() > 0.05


def test_aux_outputs_need_an_aux_frame(ctx):
    from msplat._abi import MsplatError
    from msplat.backend import DeviceGaussians, render
    from msplat.prior import aux_outputs
    sc, cam, gt, prior = prior_workload(seed=5, features=("aux",))
    render(ctx, DeviceGaussians(ctx, sc), cam, BG)
    with pytest.raises(MsplatError, match="aux"):
        aux_outputs(ctx, cam["H"], cam["W"])


def test_finalize_depth_matches_oracle(ctx, oracle):
    from msplat.prior import finalize_depth
    rng = np.random.default_rng(4)
    t = np.concatenate([[1.0, 1.0 - 1e-5, 0.5, 0.0], rng.uniform(0, 1, 4996)]).astype(np.float32)
    d = (rng.uniform(0.5, 40, 5000) * (1 - t)).astype(np.float32)
    gd, ga = finalize_depth(ctx, d, t, 0.25)
    od, oa = oracle.finalize_depth(d, t, 0.25)
    np.testing.assert_array_equal(ga, oa)
    np.testing.assert_array_equal(gd, od)
