"""The benchmark's prior-guided workload, on the CPU: the street preset's prior settings
are the reference's (STREET_PRESET, python/msplat/__init__.py:61-72, with its
exposure_compensation under PriorOptions' name exposure), and the synthetic camera
priors have the shape their docstring gives (half resolution, sky mask above the
horizon without depth, ground depth below it up to 80 m, the hood rows not kept)."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def test_street_preset_is_the_reference_preset():
    import bench
    from msplat.trainer import PriorOptions
    assert bench.STREET_PRIORS == dict(depth_weight=0.2, depth_weight_final=0.05, sky_alpha_weight=0.05,
                                       fill_weight=0.01, use_masks=True, learn_sky=True, exposure=True,
                                       max_scale_ratio=10.0)
    o = PriorOptions(**bench.STREET_PRIORS)
    assert o.active() and (o.sky_width, o.sky_height) == (512, 128)


def test_street_camera_priors_shape():
    import bench
    from msplat.scenes import street_cameras
    cams = street_cameras(4, bench.W_IMG, bench.H_IMG, bench.FOCAL, bench.FOCAL)
    for c, p in zip(cams, bench.street_camera_priors(cams)):
        ph, pw = c["H"] // 2, c["W"] // 2
        assert p["depth"].shape == (ph, pw) and p["aux"].shape == (ph, pw, 4) and p["aux"].dtype == np.uint8
        sky = p["aux"][..., 1] == 255
        assert 0.1 < sky.mean() < 0.6 and not p["depth"][sky].any()  # no depth in the sky
        assert p["depth"].max() <= 80.0 and (p["depth"] > 0).mean() > 0.4
        keep = p["aux"][..., 2] == 255
        assert keep[: int(0.9 * ph)].all() and not keep[-1].any()  # the hood rows are not kept
        assert p["has_depth"] and p["has_sky"] and p["has_mask"]
