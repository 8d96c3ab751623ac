"""The ground-truth image pipeline on the GPU (msplat_hip_resize_area, msplat_hip_undistort)
against the compiled reference's outputs (tests/golden/image_pipeline.npz), bit for bit,
and the pyramid levels against the oracle."""
import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu


def test_resize_and_undistort_match_reference(ctx):
    from msplat._abi import DeviceArray
    from msplat.images import resize_area, undistort
    from workloads import RESIZE_CASES, UNDISTORT_CASES
    d = golden("image_pipeline")
    for i, (_, sw, sh, dw, dh) in enumerate(RESIZE_CASES):
        out = resize_area(ctx, DeviceArray.from_numpy(ctx, d[f"resize_src{i}"]), dw, dh).numpy()
        np.testing.assert_array_equal(out, d[f"resize_out{i}"], err_msg=f"resize case {i}")
    src = DeviceArray.from_numpy(ctx, d["undistort_src"])
    for i, (intr, dist, rx, ry, dw, dh) in enumerate(UNDISTORT_CASES):
        out = undistort(ctx, src, intr, dist, rx, ry, dw, dh).numpy()
        np.testing.assert_array_equal(out, d[f"undistort_out{i}"], err_msg=f"undistort case {i}")


def test_target_and_pyramid_match_oracle(ctx, oracle):
    """A 1264x832 RGBA8 frame resampled to 632x416, undistorted into a crop, and its
    pyramid levels 2 and 4 (odd sizes round down, levelSize)."""
    from msplat.camera import level_size
    from msplat.images import Pyramid, load_target
    g = np.random.default_rng(12)
    rgba = g.integers(0, 256, (832, 1264, 4), dtype=np.uint8)
    intr, dist = (560.0, 560.0, 316.5, 208.2), (-0.12, 0.03, 0.001, -0.0005, 0.0)
    tgt = load_target(ctx, rgba, 632, 416, intr, dist, crop=(6, 4), crop_size=(620, 407))
    shifted = np.array(intr, np.float32) + np.array([0, 0, 6, 4], np.float32)
    ref = oracle.undistort(oracle.resize_area(rgba, 632, 416), shifted,
                           np.array(dist, np.float32), 6, 4, 620, 407)
    np.testing.assert_array_equal(tgt.numpy(), ref)
    pyr = Pyramid(ctx, tgt)
    for f in (2, 4):
        w, h = level_size(620, f), level_size(407, f)
        np.testing.assert_array_equal(pyr(f).numpy(), oracle.resize_area(ref, w, h))
    assert pyr(1) is tgt and pyr(2) is pyr(2)
