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
