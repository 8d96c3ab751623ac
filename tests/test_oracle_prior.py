"""The oracle's prior-guided path (msplat_train_step and msplat_render with a
PriorStep: aux rasterizer, learned sky, depth/sky/fill priors, exposure, needle cap;
msplat_cpu.cpp:327-402, 573-626, 787-1031, 1219-1309) against the compiled reference,
bit for bit, on the feature combinations of tests/workloads.py, and against the
committed fixture where the reference is not built."""
import numpy as np
import pytest

from conftest import golden
from workloads import BG, PRIOR_VARIANTS, prior_workload

KEYS = ("means", "scales", "quats", "dc", "rest", "opac")
PRIOR_OUT = ("sky_tex", "sky_grad", "sky_m1", "sky_m2", "loss_terms", "expo_params", "expo_m1", "expo_m2",
             "expo_grad")


def _state(sc):
    st = {k: np.ascontiguousarray(sc[k], np.float32).copy() for k in KEYS}
    for k in KEYS:
        st["m1_" + k] = np.zeros_like(st[k])
        st["m2_" + k] = np.zeros_like(st[k])
    n = st["means"].shape[0]
    st.update(degree=sc["degree"], vis=np.zeros(n, np.float32), gn=np.zeros(n, np.float32),
              ms=np.zeros(n, np.float32))
    return st


def _copy(prior):
    return {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in prior.items()}


STEPS = np.array([1.6e-4, 5e-3, 1e-3, 2.5e-3, 1.25e-4, 5e-2], np.float32)
BC2 = np.full(6, np.float32(np.sqrt(np.float32(1 - 0.999))), np.float32)


def _reference():
    from oracle import Reference, reference_available
    if not reference_available():
        pytest.skip("the compiled reference (oracle/_ref) is not built here")
    return Reference()


@pytest.mark.parametrize("variant", sorted(PRIOR_VARIANTS))
def test_prior_train_steps_equal_reference(oracle, variant):
    ref = _reference()
    sc, cam, gt, prior = prior_workload(seed=3, features=PRIOR_VARIANTS[variant])
    so, sr = _state(sc), _state(sc)
    po, pr = _copy(prior), _copy(prior)
    for t in range(2):  # two steps: the second starts from updated moments, sky and exposure
        lo, ro, _, do, fo = oracle.train_step_prior(so, cam, BG, gt, STEPS, BC2, po)
        lr, rr, dr, fr = ref.train_step_prior(sr, cam, BG, gt, STEPS, BC2, pr)
        assert lo == np.float32(lr), (t, lo, lr)
        np.testing.assert_array_equal(ro, rr)
        np.testing.assert_array_equal(fo, fr)
        if prior.get("aux"):
            np.testing.assert_array_equal(do, dr)
    for k in so:
        if k != "degree":
            np.testing.assert_array_equal(so[k], sr[k], err_msg=k)
    for k in PRIOR_OUT:
        if po.get(k) is not None:
            np.testing.assert_array_equal(po[k], pr[k], err_msg=k)
    if prior.get("aux"):
        assert np.all(np.isfinite(po["loss_terms"]))
    if "sky" in PRIOR_VARIANTS[variant]:
        assert np.any(po["sky_tex"] != prior["sky_tex"])  # the sky learned something
    if "cap" in PRIOR_VARIANTS[variant]:
        s = so["scales"]
        mid = s.sum(1) - s.max(1) - s.min(1)
        assert np.all(s.max(1) <= mid + np.float32(np.log(np.float32(3.0))) + 1e-6)


@pytest.mark.parametrize("variant", ["all", "aux_sky", "aux_defaults"])
def test_prior_render_equals_reference(oracle, variant):
    ref = _reference()
    sc, cam, gt, prior = prior_workload(seed=5, features=PRIOR_VARIANTS[variant])
    io, do, fo, _ = oracle.render_prior(sc, cam, BG, prior)
    ir, dr, fr = ref.render_prior(sc, cam, BG, prior)
    np.testing.assert_array_equal(io, ir)
    np.testing.assert_array_equal(do, dr)
    np.testing.assert_array_equal(fo, fr)


def test_prior_step_matches_fixture(oracle):
    """The oracle against the compiled reference's step on the 'all' case (fixture
    written by tests/golden/make_golden.py)."""
    d = golden("prior_step")
    sc, cam, gt, prior = prior_workload(seed=3, features=PRIOR_VARIANTS["all"])
    st = _state(sc)
    loss, radii, img, depth, fT = oracle.train_step_prior(st, cam, BG, gt, STEPS, BC2, prior)
    assert loss == np.float32(d["loss"])
    np.testing.assert_array_equal(depth, d["depth"])
    np.testing.assert_array_equal(fT, d["final_T"])
    for k in KEYS:
        np.testing.assert_array_equal(st[k], d["out_" + k].reshape(st[k].shape), err_msg=k)
    for k in ("sky_tex", "expo_params", "loss_terms"):
        np.testing.assert_array_equal(prior[k], d["out_" + k], err_msg=k)


def test_prior_closed_forms(oracle):
    """What the mathematics fixes: with every prior weight zero, no sky, no exposure and
    no cap, the prior step is the plain step bit for bit; the aux render's depth sum over
    alpha lies within the gaussians' depth range; an all-sky mask leaves only the sky term."""
    sc, cam, gt, prior = prior_workload(seed=9, features=("aux",))
    a, b = _state(sc), _state(sc)
    lp = oracle.train_step_prior(a, cam, BG, gt, STEPS, BC2, dict(prior, aux=False))[0]
    lq = oracle.train_step(b, cam, BG, gt, STEPS, BC2)[0]
    assert lp == lq
    for k in KEYS:
        np.testing.assert_array_equal(a[k], b[k])
    img, depth, fT, radii = oracle.render_prior(sc, cam, BG, prior)
    A = 1.0 - fT
    m = A > 0.5
    assert m.sum() > 20
    z = depth[m] / A[m]
    assert z.min() > 1.0 and z.max() < 7.0
    sc2, cam2, gt2, pr2 = prior_workload(seed=9, features=("aux", "sky_mask"))
    pr2["prior_aux"][..., 1] = 255
    st = _state(sc2)
    _, _, _, _, fT2 = oracle.train_step_prior(st, cam2, BG, gt2, STEPS, BC2, pr2)
    keep = pr2["prior_aux"][..., 2] >= 128
    assert keep.all()
    np.testing.assert_allclose(pr2["loss_terms"][1], 0.3 * float(np.mean(1.0 - fT2.astype(np.float64))), rtol=1e-5)
    assert pr2["loss_terms"][0] == 0 and pr2["loss_terms"][2] == 0


def test_finalize_depth_equals_reference_and_closed_form(oracle):
    """msplat_finalize_depth (cpu:1736-1745): alpha = 1 - T; depth = num / alpha * s, 0 at
    alpha <= 1e-4 (T = 1, T just below 1, T = 0.5 and T = 0 included)."""
    rng = np.random.default_rng(4)
    t = np.concatenate([[1.0, 1.0 - 1e-5, 1.0 - 2e-4, 0.5, 0.0], rng.uniform(0, 1, 995)]).astype(np.float32)
    d = (rng.uniform(0.5, 40, 1000) * (1 - t)).astype(np.float32)
    od, oa = oracle.finalize_depth(d, t, 0.25)
    np.testing.assert_array_equal(oa, (np.float32(1) - t))
    assert od[0] == 0 and od[1] == 0 and od[2] != 0
    np.testing.assert_allclose(od[3:], (d[3:].astype(np.float64) / oa[3:] * 0.25), rtol=1e-6)
    from oracle import reference_available
    if reference_available():
        from oracle import Reference
        rd, ra = Reference().finalize_depth(d, t, 0.25)
        np.testing.assert_array_equal(od, rd)
        np.testing.assert_array_equal(oa, ra)
