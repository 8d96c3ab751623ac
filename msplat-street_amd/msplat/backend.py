"""Device-resident Gaussian state and the per-view backend calls over the C ABI.

DeviceGaussians keeps the six parameter groups, their Adam moments and the
densification statistics in capacity-sized device buffers (the reference's *_buf
tensors with views of the active count, model.cpp:81-156).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from ._abi import (Adam, Context, DeviceArray, GradStats, Gaussians, Grads, Views, check, floats, lib,
                   make_camera_struct)

KEYS = ("means", "scales", "quats", "dc", "rest", "opac")
WIDTH = {"means": 3, "scales": 3, "quats": 4, "dc": 3, "opac": 1}
BETA1, BETA2, EPS = 0.9, 0.999, 1e-8


def sh_bases(degree: int) -> int:
    return {0: 1, 1: 4, 2: 9, 3: 16}.get(degree, 25)


class DeviceGaussians:
    def __init__(self, ctx: Context, scene: dict, capacity: int | None = None):
        self.ctx = ctx
        self.degree = int(scene["degree"])
        self.n = int(scene["means"].shape[0])
        self.capacity = max(int(capacity or self.n), self.n, 1)
        self.rest_bases = sh_bases(self.degree) - 1
        self.widths = dict(WIDTH, rest=max(self.rest_bases, 1) * 3)
        self.params, self.m1, self.m2 = {}, {}, {}
        for k in KEYS:
            w = self.widths[k]
            host = np.zeros((self.capacity, w), np.float32)
            if self.n:  # a degree-0 scene's rest may have no columns: its placeholder stays zero
                src = np.asarray(scene[k], np.float32).reshape(self.n, -1)[:, :w]
                host[:self.n, :src.shape[1]] = src
            self.params[k] = DeviceArray.from_numpy(ctx, host)
            self.m1[k] = DeviceArray(ctx, (self.capacity, w))
            self.m2[k] = DeviceArray(ctx, (self.capacity, w))
        self.vis = DeviceArray(ctx, (self.capacity,))
        self.gn = DeviceArray(ctx, (self.capacity,))
        self.ms = DeviceArray(ctx, (self.capacity,))
        self.radii = DeviceArray(ctx, (self.capacity,), np.int32)
        self.step_count = 0

    def arrays(self) -> list:
        """Every device array the scene owns (parameters, moments, statistics, radii)."""
        return ([self.params[k] for k in KEYS] + [self.m1[k] for k in KEYS] + [self.m2[k] for k in KEYS]
                + [self.vis, self.gn, self.ms, self.radii])

    @property
    def nbytes(self) -> int:
        """Device memory the scene holds."""
        return sum(a.nbytes for a in self.arrays() if a.ptr)

    def free(self):
        """Frees the scene's device memory now (after the work that reads it has completed)."""
        for a in self.arrays():
            a.free()

    def struct(self) -> Gaussians:
        g = Gaussians()
        g.num_points = self.n
        g.means, g.scales, g.quats = self.params["means"].ptr, self.params["scales"].ptr, self.params["quats"].ptr
        g.features_dc, g.features_rest = self.params["dc"].ptr, self.params["rest"].ptr
        g.opacities = self.params["opac"].ptr
        return g

    def adam(self, step_sizes, bc2s, beta1=BETA1, beta2=BETA2, eps=EPS) -> Adam:
        a = Adam()
        for i, k in enumerate(KEYS):
            a.exp_avg[i] = self.m1[k].ptr
            a.exp_avg_sq[i] = self.m2[k].ptr
            a.step_size[i] = float(step_sizes[i])
            a.bc2_sqrt[i] = float(bc2s[i])
        a.beta1, a.beta2, a.eps = beta1, beta2, eps
        return a

    def stats(self) -> GradStats:
        s = GradStats()
        s.vis_counts, s.xys_grad_norm, s.max_2d_size = self.vis.ptr, self.gn.ptr, self.ms.ptr
        return s

    def reset_stats(self):
        for d in (self.vis, self.gn, self.ms):
            check(lib().msplat_hip_memset(self.ctx.ptr, C.c_void_p(d.ptr), 0, d.nbytes))

    def buffers18(self):
        return [self.params[k] for k in KEYS] + [self.m1[k] for k in KEYS] + [self.m2[k] for k in KEYS]

    def ensure_capacity(self, needed: int):
        """Grow every capacity buffer to max(needed, 2 x capacity), keeping the active
        rows (Model::ensureCapacity, model.cpp:158-193)."""
        if needed <= self.capacity:
            return
        cap = max(needed, 2 * self.capacity)
        self.ctx.sync()

        def grow(old: DeviceArray) -> DeviceArray:
            new = DeviceArray(self.ctx, (cap,) + old.shape[1:], old.dtype)
            nbytes = self.n * int(np.prod(old.shape[1:], dtype=np.int64)) * old.dtype.itemsize
            if nbytes:
                check(lib().msplat_hip_memcpy(self.ctx.ptr, C.c_void_p(new.ptr), C.c_void_p(old.ptr), nbytes, 2))
            return new

        for d in (self.params, self.m1, self.m2):
            for k in KEYS:
                d[k] = grow(d[k])
        for k in ("vis", "gn", "ms", "radii"):
            setattr(self, k, grow(getattr(self, k)))
        self.ctx.sync()
        self.capacity = cap

    def densify(self, grad_thresh, size_thresh, screen_thresh, check_screen, cull_alpha, cull_scale, cull_screen,
                check_huge, half_max_dim, seed) -> int:
        bufs = (C.c_void_p * 18)(*[b.ptr for b in self.buffers18()])
        out = C.c_int32()
        check(lib().msplat_hip_densify(self.ctx.ptr, self.n, self.capacity, grad_thresh, size_thresh, screen_thresh,
                                       check_screen, cull_alpha, cull_scale, cull_screen, check_huge, self.gn.ptr,
                                       self.vis.ptr, self.ms.ptr, half_max_dim, bufs, self.widths["rest"],
                                       int(seed) & 0xFFFFFFFF, C.byref(out)))
        self.n = int(out.value)
        return self.n

    def opacity_reset(self, reset_logit: float):
        check(lib().msplat_hip_opacity_reset(self.ctx.ptr, self.params["opac"].ptr, self.m1["opac"].ptr,
                                             self.m2["opac"].ptr, self.n, reset_logit))

    @classmethod
    def from_points(cls, ctx: Context, points: np.ndarray, rgb: np.ndarray, degree: int, capacity=None, seed=42,
                    opacity_logit=None) -> "DeviceGaussians":
        """Model's initialisation (model.cpp:37-79): exact 3-NN log-scales, uniformly random
        rotations, DC from the point colours, opacity logit(0.1), zero higher bands —
        computed on the GPU (msplat_hip_init_gaussians)."""
        n = int(points.shape[0])
        nb = sh_bases(degree)
        scene = dict(means=points, scales=np.zeros((n, 3), np.float32), quats=np.zeros((n, 4), np.float32),
                     dc=np.zeros((n, 3), np.float32), rest=np.zeros((n, max(nb - 1, 1), 3), np.float32),
                     opac=np.zeros((n, 1), np.float32), degree=degree)
        dg = cls(ctx, scene, capacity)
        logit = float(np.float32(np.log(np.float32(0.1) / np.float32(0.9)))) if opacity_logit is None else opacity_logit
        drgb = DeviceArray.from_numpy(ctx, np.ascontiguousarray(rgb, np.uint8))
        check(lib().msplat_hip_init_gaussians(ctx.ptr, dg.params["means"].ptr, drgb.ptr, n, seed, logit,
                                              dg.params["scales"].ptr, dg.params["quats"].ptr, dg.params["dc"].ptr,
                                              dg.params["opac"].ptr))
        ctx.sync()
        return dg

    def to_numpy(self) -> dict:
        out = {"degree": self.degree}
        for k in KEYS:
            p = self.params[k].numpy()[:self.n]
            out[k] = p if k != "rest" else p.reshape(self.n, max(self.rest_bases, 1), 3)
            out["m1_" + k] = self.m1[k].numpy()[:self.n]
            out["m2_" + k] = self.m2[k].numpy()[:self.n]
        for k in ("vis", "gn", "ms"):
            out[k] = getattr(self, k).numpy()[:self.n]
        return out


def render(ctx: Context, dg: DeviceGaussians, cam: dict, bg) -> int:
    """Enqueues a render; returns the device pointer of the (H, W, 3) image."""
    g, c = dg.struct(), make_camera_struct(cam, dg.degree)
    out = C.c_void_p()
    check(lib().msplat_hip_render(ctx.ptr, C.byref(g), C.byref(c), floats(bg), C.byref(out)))
    return out.value


def render_numpy(ctx: Context, dg: DeviceGaussians, cam: dict, bg) -> np.ndarray:
    p = render(ctx, dg, cam, bg)
    return ctx.download(p, (cam["H"], cam["W"], 3), np.float32)


def adam_coeffs(lrs, step_count, beta1=BETA1, beta2=BETA2):
    """(step sizes lr/bc1, bc2_sqrt) for Adam step t (model.cpp:638-651): the reference's
    std::pow(float beta, int t) promotes the float beta to double, and 1 - beta^t is
    rounded to float once; the step sizes and square root are float."""
    t = int(step_count)
    bc1 = np.float32(1.0 - float(np.float32(beta1)) ** t)
    bc2 = np.float32(1.0 - float(np.float32(beta2)) ** t)
    steps = (np.asarray(lrs, np.float32) / bc1).astype(np.float32)
    return steps, np.full(6, np.sqrt(bc2), np.float32)


def train_step(ctx: Context, dg: DeviceGaussians, cam: dict, bg, gt: DeviceArray, step_sizes, bc2s,
               ssim_weight=0.2, read_loss=True, beta1=BETA1, beta2=BETA2, eps=EPS) -> float | None:
    g, c = dg.struct(), make_camera_struct(cam, dg.degree)
    a, s = dg.adam(step_sizes, bc2s, beta1, beta2, eps), dg.stats()
    H, W = cam["H"], cam["W"]
    inv_n = np.float32(1.0) / np.float32(H * W * 3)
    inv_max = np.float32(1.0) / np.float32(max(H, W))
    loss = C.c_float(0.0)
    check(lib().msplat_hip_train_step(ctx.ptr, C.byref(g), C.byref(c), floats(bg), C.c_void_p(gt.ptr), ssim_weight,
                                      float(inv_n), C.byref(a), C.byref(s), float(inv_max),
                                      C.c_void_p(dg.radii.ptr), C.byref(loss) if read_loss else None))
    return loss.value if read_loss else None


class BatchViews:
    """The C views record of a batch (cameras, backgrounds, optional target pointers);
    keeps the ctypes arrays alive as long as it lives."""

    def __init__(self, cams, degree: int, bgs, targets=None):
        n = len(cams)
        self.cams = list(cams)
        self.c_cams = (_abi.Camera * n)(*[make_camera_struct(c, degree) for c in cams])
        bg = np.asarray(bgs, np.float32).reshape(-1, 3)
        if bg.shape[0] == 1:
            bg = np.repeat(bg, n, 0)
        self.c_bgs = (C.c_float * (3 * n))(*bg.reshape(-1).tolist())
        # the record holds raw device pointers: keep the arrays alive as long as it lives
        self.targets = list(targets) if targets is not None else None
        self.c_targets = (C.c_void_p * n)(*[t.ptr if isinstance(t, DeviceArray) else t for t in targets]) \
            if targets is not None else None
        self.rec = Views()
        self.rec.num_views = n
        self.rec.cameras = self.c_cams
        self.rec.backgrounds = self.c_bgs
        self.rec.targets = self.c_targets if self.c_targets is not None else None
        self.sizes = [(int(c["H"]), int(c["W"])) for c in cams]

    def offsets(self):
        out, acc = [], 0
        for h, w in self.sizes:
            out.append(acc)
            acc += h * w
        return out, acc


def render_batch(ctx: Context, dg: DeviceGaussians, views: BatchViews) -> tuple:
    """Renders every view (syncs); returns (device pointer of the packed images, pixel offsets)."""
    g = dg.struct()
    out = C.c_void_p()
    offs = (C.c_uint64 * len(views.sizes))()
    check(lib().msplat_hip_render_batch(ctx.ptr, C.byref(g), C.byref(views.rec), C.byref(out), offs))
    return out.value, list(offs)


def render_batch_numpy(ctx: Context, dg: DeviceGaussians, views: BatchViews) -> list:
    ptr, offs = render_batch(ctx, dg, views)
    _, total = views.offsets()
    packed = ctx.download(ptr, (total, 3), np.float32)
    return [packed[o:o + h * w].reshape(h, w, 3) for o, (h, w) in zip(offs, views.sizes)]


def train_step_batch(ctx: Context, dg: DeviceGaussians, views: BatchViews, step_sizes, bc2s, ssim_weight=0.2,
                     beta1=BETA1, beta2=BETA2, eps=EPS) -> np.ndarray:
    """One batched step over every view (syncs); returns the per-view losses."""
    g = dg.struct()
    a, s = dg.adam(step_sizes, bc2s, beta1, beta2, eps), dg.stats()
    losses = (C.c_float * len(views.sizes))()
    check(lib().msplat_hip_train_step_batch(ctx.ptr, C.byref(g), C.byref(views.rec), ssim_weight, C.byref(a),
                                            C.byref(s), losses))
    return np.array(losses[:], np.float32)


def render_backward_batch(ctx: Context, dg: DeviceGaussians, views: BatchViews, v_images: DeviceArray) -> dict:
    """Gradients of sum_v <image_v, v_image_v> for the last render_batch, summed over the views."""
    n = dg.n
    out = {k: DeviceArray(ctx, (max(n, 1), dg.widths[k])) for k in KEYS}
    gr = Grads(*[out[k].ptr for k in KEYS])
    check(lib().msplat_hip_render_backward_batch(ctx.ptr, C.byref(dg.struct()), C.byref(views.rec),
                                                 C.c_void_p(v_images.ptr), C.byref(gr)))
    return {k: out[k].numpy()[:n] for k in KEYS}


def render_backward(ctx: Context, dg: DeviceGaussians, cam: dict, bg, v_img: DeviceArray) -> dict:
    """Gradients of <image, v_img> for the last render of `cam` on this context."""
    n = dg.n
    out = {k: DeviceArray(ctx, (max(n, 1), dg.widths[k])) for k in KEYS}
    gr = Grads(*[out[k].ptr for k in KEYS])
    g, c = dg.struct(), make_camera_struct(cam, dg.degree)
    check(lib().msplat_hip_render_backward(ctx.ptr, C.byref(g), C.byref(c), floats(bg), C.c_void_p(v_img.ptr),
                                           C.byref(gr)))
    return {k: out[k].numpy()[:n] for k in KEYS}
