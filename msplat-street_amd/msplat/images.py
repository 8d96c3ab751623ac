"""The ground-truth image pipeline on the GPU: a decoded RGBA8 image becomes the float RGB
training target at the working resolution (area resampling, then undistortion into the
crop when the camera has lens distortion), and the downscale pyramid levels come from
that target (Camera::loadImage and Camera::getGPUImage, core/src/input_data.cpp:34-70,
over msplat_resize_area and msplat_undistort).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._abi import Context, DeviceArray, check, floats, lib
from .camera import level_size


def resize_area(ctx: Context, src: DeviceArray, dw: int, dh: int) -> DeviceArray:
    """(sh, sw, 4) uint8 RGBA or (sh, sw, 3) float32 on the device -> (dh, dw, 3) float32."""
    sh, sw = src.shape[:2]
    is_u8 = src.dtype == np.uint8
    if is_u8 and src.shape[2] != 4 or not is_u8 and (src.dtype != np.float32 or src.shape[2] != 3):
        raise ValueError("resize_area takes (h, w, 4) uint8 or (h, w, 3) float32")
    out = DeviceArray(ctx, (dh, dw, 3), np.float32, zero=False)
    check(lib().msplat_hip_resize_area(ctx.ptr, C.c_void_p(src.ptr), int(is_u8), sw, sh, C.c_void_p(out.ptr), dw, dh))
    return out


def undistort(ctx: Context, src: DeviceArray, intr, dist, roi_x: int, roi_y: int, dw: int, dh: int) -> DeviceArray:
    """Brown-Conrady undistortion of (sh, sw, 3) float32 into the (dh, dw) crop at (roi_x, roi_y);
    intr = (fx, fy, cx, cy), dist = (k1, k2, p1, p2, k3)."""
    sh, sw = src.shape[:2]
    out = DeviceArray(ctx, (dh, dw, 3), np.float32, zero=False)
    check(lib().msplat_hip_undistort(ctx.ptr, C.c_void_p(src.ptr), sw, sh, floats(intr), floats(dist), roi_x, roi_y,
                                     C.c_void_p(out.ptr), dw, dh))
    return out


def load_target(ctx: Context, rgba: np.ndarray, width: int, height: int, intr=None, dist=None, crop=(0, 0),
                crop_size=None) -> DeviceArray:
    """A decoded (h, w, 4) RGBA8 image -> the float training target: area-resampled to
    (width, height); with `dist`, undistorted into the crop of size crop_size at `crop`,
    using the resampled image's intrinsics `intr` (input_data.cpp:34-46)."""
    src = DeviceArray.from_numpy(ctx, np.ascontiguousarray(rgba, np.uint8))
    level = resize_area(ctx, src, width, height)
    if dist is not None and any(float(v) != 0.0 for v in dist):
        cw, ch = crop_size if crop_size else (width, height)
        fx, fy, cx, cy = (np.float32(v) for v in intr)  # the crop shifts the principal point, in float
        out = undistort(ctx, level, (fx, fy, cx + np.float32(crop[0]), cy + np.float32(crop[1])), dist, crop[0],
                        crop[1], cw, ch)
        ctx.sync()
        level.free()
        level = out
    ctx.sync()
    src.free()
    return level


class Pyramid:
    """Downscaled copies of a target, made on first use and kept (Camera::getGPUImage,
    input_data.cpp:56-70): level f has levelSize(W, f) x levelSize(H, f) pixels."""

    def __init__(self, ctx: Context, image: DeviceArray):
        self.ctx, self.image, self.levels = ctx, image, {}

    def __call__(self, factor: int) -> DeviceArray:
        if factor <= 1:
            return self.image
        if factor not in self.levels:
            h, w = self.image.shape[:2]
            self.levels[factor] = resize_area(self.ctx, self.image, level_size(w, factor), level_size(h, factor))
        return self.levels[factor]
