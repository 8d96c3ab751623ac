"""Prior-guided training on the HIP backend (PriorStep, core/metal/bindings.h:41-77):
the aux rasterizer with the learned sky, depth / sky / fill priors, per-image exposure
and the needle cap, through msplat_hip_train_step_prior / msplat_hip_render_prior.

A prior is described by a dict with PriorStep's field names (aux, sky, sky_frame,
sky_step, sky_bc2, has_depth, has_sky_mask, depth_weight, sky_weight, fill_weight,
huber_delta, min_alpha, mask_photometric, exposure, cam_index, expo_step, expo_bc2,
expo_reg, log_max_scale_ratio) and numpy arrays for its tensors (sky_tex, sky_grad,
sky_m1, sky_m2 (h, w, 3) float32; prior_depth (ph, pw) float32; prior_aux (ph, pw, 4)
uint8; loss_terms [4]; expo_params, expo_m1, expo_m2 (cams, 12); expo_grad [12]).
`DevicePrior` keeps those tensors in device memory across steps."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._abi import Context, DeviceArray, Prior, check, floats, lib, make_camera_struct
from .backend import DeviceGaussians

ARRAYS = {"sky_tex": np.float32, "sky_grad": np.float32, "sky_m1": np.float32, "sky_m2": np.float32,
          "prior_depth": np.float32, "prior_aux": np.uint8, "loss_terms": np.float32, "expo_params": np.float32,
          "expo_m1": np.float32, "expo_m2": np.float32, "expo_grad": np.float32}
INTS = ("aux", "sky", "has_depth", "has_sky_mask", "mask_photometric", "exposure", "cam_index")
FLOATS = {"sky_step": "sky_step_size", "sky_bc2": "sky_bc2_sqrt", "depth_weight": "depth_weight",
          "sky_weight": "sky_weight", "fill_weight": "fill_weight", "huber_delta": "huber_delta",
          "min_alpha": "min_alpha", "expo_step": "expo_step_size", "expo_bc2": "expo_bc2_sqrt",
          "expo_reg": "expo_reg", "log_max_scale_ratio": "log_max_scale_ratio"}
DEFAULTS = dict(huber_delta=0.05, min_alpha=0.25, mask_photometric=True, sky_bc2=1.0, expo_bc2=1.0)
POINTERS = {"sky_tex": "sky_tex", "sky_grad": "sky_grad", "sky_m1": "sky_exp_avg", "sky_m2": "sky_exp_avg_sq",
            "prior_depth": "prior_depth", "prior_aux": "prior_aux", "loss_terms": "loss_terms",
            "expo_params": "expo_params", "expo_m1": "expo_exp_avg", "expo_m2": "expo_exp_avg_sq",
            "expo_grad": "expo_grad"}


class DevicePrior:
    """A prior's tensors in device memory and its scalar fields (update them between
    steps through `fields`, e.g. the per-step Adam scalars and the camera index)."""

    def __init__(self, ctx: Context, prior: dict):
        self.ctx = ctx
        self.fields = {k: v for k, v in prior.items() if k not in ARRAYS}
        self.arrays = {}
        for k, dt in ARRAYS.items():
            a = prior.get(k)
            if a is not None:
                self.arrays[k] = DeviceArray.from_numpy(ctx, np.ascontiguousarray(a, dt), dt)
        self.shapes = {k: np.asarray(prior[k]).shape for k in self.arrays}

    def struct(self) -> Prior:
        st = Prior()
        f = dict(DEFAULTS)
        f.update(self.fields)
        for k in INTS:
            setattr(st, k, int(f.get(k, 0)))
        for k, name in FLOATS.items():
            setattr(st, name, float(f.get(k, 0.0)))
        st.sky_frame[:] = [float(x) for x in f.get("sky_frame", (0.0,) * 9)]
        for k, a in self.arrays.items():
            setattr(st, POINTERS[k], a.ptr)
        if "sky_tex" in self.shapes:
            st.sky_h, st.sky_w = self.shapes["sky_tex"][:2]
        if "prior_depth" in self.shapes:
            st.prior_h, st.prior_w = self.shapes["prior_depth"][:2]
        if "expo_params" in self.shapes:
            st.num_cams = self.shapes["expo_params"][0]
        return st

    def numpy(self, key: str) -> np.ndarray:
        return self.arrays[key].numpy().reshape(self.shapes[key])


def train_step_prior(ctx: Context, dg: DeviceGaussians, cam: dict, bg, gt: DeviceArray, step_sizes, bc2s,
                     prior: DevicePrior, ssim_weight=0.2, read_loss=True, beta1=0.9, beta2=0.999, eps=1e-8):
    """One msplat_hip_train_step_prior; returns the photometric loss when read_loss."""
    g, c = dg.struct(), make_camera_struct(cam, dg.degree)
    a, s = dg.adam(step_sizes, bc2s, beta1, beta2, eps), dg.stats()
    H, W = cam["H"], cam["W"]
    inv_n = np.float32(1.0) / np.float32(H * W * 3)
    inv_max = np.float32(1.0) / np.float32(max(H, W))
    loss = C.c_float(0.0)
    p = prior.struct()
    check(lib().msplat_hip_train_step_prior(ctx.ptr, C.byref(g), C.byref(c), floats(bg), C.c_void_p(gt.ptr),
                                            ssim_weight, float(inv_n), C.byref(a), C.byref(s), float(inv_max),
                                            C.c_void_p(dg.radii.ptr), C.byref(loss) if read_loss else None,
                                            C.byref(p)))
    return loss.value if read_loss else None


def render_prior(ctx: Context, dg: DeviceGaussians, cam: dict, bg, prior: DevicePrior) -> np.ndarray:
    """msplat_hip_render_prior: the (H, W, 3) image."""
    g, c = dg.struct(), make_camera_struct(cam, dg.degree)
    out = C.c_void_p()
    p = prior.struct()
    check(lib().msplat_hip_render_prior(ctx.ptr, C.byref(g), C.byref(c), floats(bg), C.byref(p), C.byref(out)))
    return ctx.download(out.value, (cam["H"], cam["W"], 3), np.float32)


def aux_outputs(ctx: Context, H: int, W: int) -> tuple:
    """msplat_hip_render_aux_outputs: (expected-depth sum, final transmittance), (H, W)."""
    d, t = C.c_void_p(), C.c_void_p()
    check(lib().msplat_hip_render_aux_outputs(ctx.ptr, C.byref(d), C.byref(t)))
    return ctx.download(d.value, (H, W), np.float32), ctx.download(t.value, (H, W), np.float32)


def finalize_depth(ctx: Context, depth_num, final_T, inv_scale: float = 1.0) -> tuple:
    """msplat_hip_finalize_depth on host arrays: (expected depth, alpha)."""
    d = DeviceArray.from_numpy(ctx, np.ascontiguousarray(depth_num, np.float32))
    t = DeviceArray.from_numpy(ctx, np.ascontiguousarray(final_T, np.float32))
    od, oa = DeviceArray(ctx, d.shape), DeviceArray(ctx, d.shape)
    check(lib().msplat_hip_finalize_depth(ctx.ptr, C.c_void_p(d.ptr), C.c_void_p(t.ptr), int(np.size(depth_num)),
                                          float(inv_scale), C.c_void_p(od.ptr), C.c_void_p(oa.ptr)))
    return od.numpy(), oa.numpy()
