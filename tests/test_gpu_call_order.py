"""Stage entry points called out of order are refused on the host with MSPLAT_ERR_ARG,
before any kernel launch: a raster or loss stage with no binned frame, a backward with no
render, and any of them after msplat_hip_radix_sort reused the frame's key buffers. Each
refusal leaves the context usable: a full render afterwards still matches a fresh one.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ERR_ARG = 1
BG = np.array([0.613, 0.0101, 0.3984], np.float32)


def _scene():
    from msplat.camera import look_at, make_camera
    from msplat.scenes import random_scene
    sc = random_scene(3, 500, degree=1, log_scale=(-3.5, 0.5))
    cam = make_camera(look_at((2.5, 0.7, 2.9), (0.1, -0.2, 0)), 64, 48, 60.0, 60.0, dtu=1)
    return sc, cam


def test_out_of_order_stages_are_refused():
    from msplat._abi import Context, DeviceArray, Grads, check, floats, lib, make_camera_struct
    from msplat.backend import DeviceGaussians, render_numpy
    sc, cam = _scene()
    ctx = Context(0)
    try:
        dg = DeviceGaussians(ctx, sc)
        g, c = dg.struct(), make_camera_struct(cam, 1)
        L = lib()
        gt = DeviceArray(ctx, (cam["H"], cam["W"], 3))
        v_img = DeviceArray(ctx, (cam["H"], cam["W"], 3))
        out = {k: DeviceArray(ctx, (dg.n, dg.widths[k])) for k in ("means", "scales", "quats", "dc", "rest", "opac")}
        gr = Grads(*[out[k].ptr for k in ("means", "scales", "quats", "dc", "rest", "opac")])
        loss = C.c_float()
        # a fresh context holds no frame
        assert L.msplat_hip_stage_bin(ctx.ptr, C.byref(c)) == ERR_ARG
        assert L.msplat_hip_stage_raster_forward(ctx.ptr, C.byref(c), floats(BG)) == ERR_ARG
        assert L.msplat_hip_stage_loss(ctx.ptr, C.byref(c), C.c_void_p(gt.ptr), 0.2, 1e-3, C.byref(loss)) == ERR_ARG
        assert L.msplat_hip_render_backward(ctx.ptr, C.byref(g), C.byref(c), floats(BG), C.c_void_p(v_img.ptr),
                                            C.byref(gr)) == ERR_ARG
        # projected but not binned
        check(L.msplat_hip_stage_project(ctx.ptr, C.byref(g), C.byref(c)))
        assert L.msplat_hip_stage_raster_forward(ctx.ptr, C.byref(c), floats(BG)) == ERR_ARG
        assert L.msplat_hip_render_backward(ctx.ptr, C.byref(g), C.byref(c), floats(BG), C.c_void_p(v_img.ptr),
                                            C.byref(gr)) == ERR_ARG
        # binned, not rasterized: the loss and the backward still need an image
        check(L.msplat_hip_stage_bin(ctx.ptr, C.byref(c)))
        assert L.msplat_hip_stage_loss(ctx.ptr, C.byref(c), C.c_void_p(gt.ptr), 0.2, 1e-3, C.byref(loss)) == ERR_ARG
        # a rendered frame, then a radix sort that reuses its key buffers
        ref = render_numpy(ctx, dg, cam, BG)
        keys = DeviceArray.from_numpy(ctx, np.arange(100, 0, -1, dtype=np.uint64))
        vals = DeviceArray.from_numpy(ctx, np.arange(100, dtype=np.int32))
        check(L.msplat_hip_radix_sort(ctx.ptr, C.c_void_p(keys.ptr), C.c_void_p(vals.ptr), 100, 32))
        assert L.msplat_hip_stage_raster_forward(ctx.ptr, C.byref(c), floats(BG)) == ERR_ARG
        assert L.msplat_hip_render_backward(ctx.ptr, C.byref(g), C.byref(c), floats(BG), C.c_void_p(v_img.ptr),
                                            C.byref(gr)) == ERR_ARG
        assert L.msplat_hip_stage_loss(ctx.ptr, C.byref(c), C.c_void_p(gt.ptr), 0.2, 1e-3, C.byref(loss)) == ERR_ARG
        np.testing.assert_array_equal(keys.numpy(), np.arange(1, 101, dtype=np.uint64))
        # re-binning the still-projected frame makes it usable again
        check(L.msplat_hip_stage_bin(ctx.ptr, C.byref(c)))
        check(L.msplat_hip_stage_raster_forward(ctx.ptr, C.byref(c), floats(BG)))
        img = ctx.download(ctx.frame().image, (cam["H"], cam["W"], 3), np.float32)
        np.testing.assert_array_equal(img, ref)
        np.testing.assert_array_equal(render_numpy(ctx, dg, cam, BG), ref)
        ctx.sync()
    finally:
        ctx.close()
