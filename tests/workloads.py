"""Seeded workloads shared by the fixture generator (tests/golden/make_golden.py) and the
tests that replay them."""
import numpy as np

from msplat.scenes import random_scene, ring_cameras

BG = np.array([0.613, 0.0101, 0.3984], np.float32)
C1_SCENE_SEED, C1_GT_SEED, C1_STEPS, C1_SAMPLE = 1000, 1001, 300, 256
C1_LRS = (1.6e-4, 5e-3, 1e-3, 2.5e-3, 1.25e-4, 5e-2)


def c1_workload():
    """The CPU-runnable plain workload: 20,000 Gaussians (means U[-1,1]^3, log-scales
    N(-4.5, 0.5), uniform rotations, logits N(0,1), SH degree 1 with DC N(0, 0.5) and the
    band N(0, 0.1)); targets rendered from a second seeded 20,000-Gaussian set; 8 pinhole
    cameras at 128x128, fx = fy = 120, on a radius-4 ring looking at the origin."""
    sc = random_scene(C1_SCENE_SEED, 20000, degree=1)
    gt_scene = random_scene(C1_GT_SEED, 20000, degree=1)
    return sc, gt_scene, ring_cameras(8, 4.0, 128, 128, 120.0, 120.0, dtu=1)


def c1_camera_order(steps, count=8, seed=42):
    """The trainer's view order: a fresh permutation of the views per epoch from one
    seeded generator."""
    rng = np.random.default_rng(seed)
    order = []
    while len(order) < steps:
        order.extend(int(i) for i in rng.permutation(count))
    return np.array(order[:steps], np.int32)


# image-pipeline fixtures: (source kind, sw, sh, dw, dh); (intrinsics, distortion, roi x, roi y, dw, dh)
RESIZE_CASES = [("u8", 97, 61, 40, 25), ("u8", 40, 30, 40, 30), ("u8", 31, 17, 62, 40), ("f32", 64, 48, 21, 13),
                ("f32", 33, 20, 33, 20), ("f32", 1, 1, 3, 2), ("f32", 96, 64, 48, 32)]
UNDISTORT_CASES = [((70.0, 72.0, 40.3, 29.7), (0.1, -0.05, 0.001, -0.002, 0.01), 3, 2, 70, 50),
                   ((70.0, 72.0, 40.3, 29.7), (-0.3, 0.12, -0.004, 0.003, -0.02), -4, -3, 90, 70),
                   ((60.0, 60.0, 40.0, 30.0), (0.0, 0.0, 0.0, 0.0, 0.0), 0, 0, 80, 60)]


def prior_workload(seed=0, n=600, W=64, H=48, features=("aux", "sky", "depth", "sky_mask", "fill", "exposure", "cap",
                                                         "unkept"), sky=(16, 32)):
    """A prior-guided training case after the reference's gradcheck priors run: a few
    hundred SH1 gaussians in front of a W x H camera (so that sky shows between them), a
    target, and a PriorStep dict with the named features: the aux rasterizer, a learned
    sky (`sky` = (rows, columns), 32x16 by default), a depth prior at its own resolution (W/2 + 3 by H/2 - 1, with invalid
    samples), a sky mask over its top quarter, the fill term, exposure rows for 2 cameras,
    the needle cap, and pixels the prior does not keep. Every array is freshly made."""
    from msplat.scenes import random_scene, ring_cameras
    rng = np.random.default_rng(seed + 7)
    sc = random_scene(seed, n, degree=1, log_scale=(-3.0, 0.4))
    cam = ring_cameras(1, 3.5, W, H, 0.9 * W, 0.95 * W, dtu=1)[0]
    gt = rng.uniform(0, 1, (H, W, 3)).astype(np.float32)
    pw, ph = W // 2 + 3, H // 2 - 1
    yy, xx = np.mgrid[0:ph, 0:pw]
    depth = (3.2 * (1.0 + 0.25 * np.sin(0.3 * xx) * np.cos(0.2 * yy))).astype(np.float32)
    depth[(xx * 7 + yy * 3) % 23 == 0] = 0.0
    aux = np.zeros((ph, pw, 4), np.uint8)
    aux[..., 0] = 60 + (xx * 13 + yy * 7) % 190
    aux[..., 1] = np.where(yy < ph // 4, 255, 0)
    aux[..., 2] = 255
    if "unkept" in features:
        aux[..., 2] = np.where((xx + 2 * yy) % 11 == 0, 0, 255)
    f = set(features)
    prior = dict(aux="aux" in f, loss_terms=np.zeros(4, np.float32), huber_delta=0.05, min_alpha=0.25,
                 mask_photometric=True)
    if "sky" in f:
        sh_, sw_ = sky
        prior.update(sky=True, sky_tex=rng.uniform(0.2, 0.8, (sh_, sw_, 3)).astype(np.float32),
                     sky_grad=np.zeros((sh_, sw_, 3), np.float32), sky_m1=np.zeros((sh_, sw_, 3), np.float32),
                     sky_m2=np.zeros((sh_, sw_, 3), np.float32), sky_frame=(0, -1, 0, 0, 0, 1, 1, 0, 0),
                     sky_step=0.01, sky_bc2=float(np.sqrt(np.float32(1 - 0.999))))
    if f & {"depth", "sky_mask", "fill"}:
        prior.update(prior_depth=depth, prior_aux=aux, has_depth="depth" in f, has_sky_mask="sky_mask" in f,
                     depth_weight=0.5 if "depth" in f else 0.0, sky_weight=0.3, fill_weight=0.1 if "fill" in f else 0.0)
    if "exposure" in f:
        ident = np.array([1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0], np.float32)
        prior.update(exposure=True, cam_index=1,
                     expo_params=(ident + rng.uniform(-0.05, 0.05, (2, 12))).astype(np.float32),
                     expo_m1=np.zeros((2, 12), np.float32), expo_m2=np.zeros((2, 12), np.float32),
                     expo_grad=rng.uniform(-1e-3, 1e-3, 12).astype(np.float32), expo_step=0.002,
                     expo_bc2=float(np.sqrt(np.float32(1 - 0.999))), expo_reg=0.01)
    if "cap" in f:
        prior["log_max_scale_ratio"] = float(np.log(np.float32(3.0)))
    return sc, cam, gt, prior


PRIOR_VARIANTS = {
    "all": ("aux", "sky", "depth", "sky_mask", "fill", "exposure", "cap", "unkept"),
    "aux_defaults": ("aux",),
    "aux_sky": ("aux", "sky"),
    "depth_only": ("aux", "depth"),
    "exposure_only": ("exposure",),
    "cap_only": ("cap",),
}
