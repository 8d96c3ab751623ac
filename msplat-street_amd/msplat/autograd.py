"""Differentiable rendering for PyTorch: a torch.autograd.Function over the HIP backend.

    img = rasterize(means, scales, quats, features_dc, features_rest, opacities, camera, bg)

Inputs are float32 CUDA(HIP) tensors in the reference layouts (means (N,3), log-scales
(N,3), quaternions (N,4) stored (w,x,y,z), features_dc (N,3), features_rest (N,B-1,3),
opacity logits (N,1) or (N,)); `camera` is a dict from msplat.camera.make_camera. The
output is the (H, W, 3) image of msplat_render (core/metal/bindings.h:81-91). Backward
returns gradients for every parameter (msplat_hip_render_backward); as in the reference
(core/cpu/msplat_cpu.cpp:1111-1137) the SH view direction is not differentiated.

The work runs on torch's current stream, so it orders with surrounding torch work.
"""
from __future__ import annotations

import ctypes as C

import torch

from ._abi import Context, Gaussians, Grads, check, floats, lib, make_camera_struct

_contexts: dict = {}


def _context(device: torch.device) -> Context:
    idx = device.index if device.index is not None else torch.cuda.current_device()
    ctx = _contexts.get(idx)
    if ctx is None:
        ctx = _contexts[idx] = Context(idx)
    stream = torch.cuda.current_stream(device).cuda_stream
    check(lib().msplat_hip_ctx_set_stream(ctx.ptr, C.c_void_p(stream)))
    return ctx


def _degree(rest: torch.Tensor) -> int:
    b = rest.shape[1] + 1 if rest.dim() == 3 else 1
    return 0 if b <= 1 else 1 if b <= 4 else 2 if b <= 9 else 3 if b <= 16 else 4


def _struct(means, scales, quats, dc, rest, opac) -> Gaussians:
    g = Gaussians()
    g.num_points = means.shape[0]
    g.means, g.scales, g.quats = means.data_ptr(), scales.data_ptr(), quats.data_ptr()
    g.features_dc, g.features_rest, g.opacities = dc.data_ptr(), rest.data_ptr(), opac.data_ptr()
    return g


def _check_inputs(tensors):
    for t in tensors:
        if not t.is_cuda or t.dtype != torch.float32:
            raise TypeError("msplat rasterize expects float32 device tensors")


class RasterizeGaussians(torch.autograd.Function):
    @staticmethod
    def forward(fctx, means, scales, quats, features_dc, features_rest, opacities, camera, background):
        ts = [t.contiguous() for t in (means, scales, quats, features_dc, features_rest, opacities)]
        _check_inputs(ts)
        ctx = _context(ts[0].device)
        deg = _degree(ts[4])
        cam = make_camera_struct(camera, deg)
        bg = floats(background)
        out = C.c_void_p()
        check(lib().msplat_hip_render(ctx.ptr, C.byref(_struct(*ts)), C.byref(cam), bg, C.byref(out)))
        H, W = camera["H"], camera["W"]
        img = torch.empty((H, W, 3), dtype=torch.float32, device=ts[0].device)
        check(lib().msplat_hip_memcpy(ctx.ptr, C.c_void_p(img.data_ptr()), out, img.numel() * 4, 2))
        fctx.save_for_backward(*ts)
        fctx.camera, fctx.background = camera, tuple(float(v) for v in background)
        return img

    @staticmethod
    def backward(fctx, v_img):
        ts = fctx.saved_tensors
        ctx = _context(ts[0].device)
        deg = _degree(ts[4])
        cam = make_camera_struct(fctx.camera, deg)
        bg = floats(fctx.background)
        g = _struct(*ts)
        # the context's frame cache may hold a later render: rebuild this one first
        out = C.c_void_p()
        check(lib().msplat_hip_render(ctx.ptr, C.byref(g), C.byref(cam), bg, C.byref(out)))
        v_img = v_img.contiguous().float()
        grads = [torch.zeros_like(t) for t in ts]
        gr = Grads(*[x.data_ptr() for x in grads])
        check(lib().msplat_hip_render_backward(ctx.ptr, C.byref(g), C.byref(cam), bg, C.c_void_p(v_img.data_ptr()),
                                               C.byref(gr)))
        return (*grads, None, None)


def rasterize(means, scales, quats, features_dc, features_rest, opacities, camera: dict, background=(0.0, 0.0, 0.0)):
    """(H, W, 3) image; differentiable in every parameter tensor."""
    return RasterizeGaussians.apply(means, scales, quats, features_dc, features_rest, opacities, camera, background)
