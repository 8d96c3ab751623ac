"""Training loop over the HIP backend: the schedule of msplat's Model
(core/src/model.cpp) — Adam bias corrections (:638-651), the means learning-rate decay
applied after each step (:140-143), the SH degree ramp (:535), densify/prune and opacity
reset cadence (:199-245) with capacity growth (:158-193) — driving one
msplat_hip_train_step per view.
"""
from __future__ import annotations

import ctypes as C
import dataclasses

import numpy as np

from ._abi import Context, DeviceArray, Prior, check, floats, lib, make_camera_struct
from .backend import KEYS, DeviceGaussians, adam_coeffs
from .camera import expf, logf

F = np.float32
LR_INIT = (1.6e-4, 5e-3, 1e-3, 2.5e-3, 1.25e-4, 5e-2)  # model.cpp:100
MEANS_LR_FINAL = 1.6e-6
RESET_LOGIT = -1.3862943611198906  # logit(0.2)

# afterTrain's decisions at one step (model.cpp:199-245)
REFINE_DENSIFY, REFINE_RESET, REFINE_STATS, REFINE_SCREEN, REFINE_HUGE = 1, 2, 4, 8, 16


def means_lr(step: int, max_steps: int, lr_init=LR_INIT[0], lr_final=MEANS_LR_FINAL) -> float:
    """Model::schedulersStep (model.cpp:140-143) in float32 through the C library's
    expf/logf: exp(log(init) (1 - t) + log(final) t), t = clamp(step / max_steps, 0, 1)."""
    t = F(min(max(F(step) / F(max_steps), F(0)), F(1)))
    return float(expf(logf(lr_init) * (F(1) - t) + logf(lr_final) * t))


def refine_flags(step: int, cfg: "TrainConfig", num_cameras: int, stop_split_at: int) -> int:
    """What Model::afterTrain does after `step` (model.cpp:199-245), as REFINE_* bits."""
    if step % cfg.refine_every != 0 or step <= cfg.warmup_length:
        return 0
    reset_interval = cfg.reset_alpha_every * cfg.refine_every
    f = REFINE_STATS
    if step < stop_split_at and step % reset_interval > num_cameras + cfg.refine_every:
        f |= REFINE_DENSIFY
        if step < cfg.stop_screen_size_at:
            f |= REFINE_SCREEN
        if step > cfg.refine_every * cfg.reset_alpha_every:
            f |= REFINE_HUGE
    if step < stop_split_at and step % reset_interval == cfg.refine_every:
        f |= REFINE_RESET
    return f


@dataclasses.dataclass
class TrainConfig:
    """The TrainingConfig fields of python/bindings.cpp:27-55 that the plain method uses,
    plus the cull threshold that Model hard-codes to 0.1 (model.cpp:218)."""
    iterations: int = 30000
    sh_degree: int = 3
    sh_degree_interval: int = 1000
    ssim_weight: float = 0.2
    num_downscales: int = 0         # the reference's default is 2 (bindings.cpp:494); 0 trains at full size
    resolution_schedule: int = 3000
    refine_every: int = 100
    warmup_length: int = 500
    reset_alpha_every: int = 30
    densify_grad_thresh: float = 0.0002
    densify_size_thresh: float = 0.01
    stop_screen_size_at: int = 4000
    split_screen_size: float = 0.05
    cull_alpha_thresh: float = 0.1
    cull_scale_thresh: float = 0.5
    cull_screen_size: float = 0.15
    bg_color: tuple = (0.6130, 0.0101, 0.3984)
    densify: bool = True
    seed: int = 42
    step_graph: bool = False        # replay each step from a captured hipGraph (msplat_hip_set_step_graph)


@dataclasses.dataclass
class PriorOptions:
    """PriorOptions (core/include/model.hpp:18-33): the prior-guided training settings."""
    depth_weight: float = 0.0         # log-depth prior weight at the first step (0: off)
    depth_weight_final: float = -1.0  # weight at the last step, log-linear in between (< 0: constant)
    depth_huber_delta: float = 0.05
    depth_min_alpha: float = 0.25
    sky_alpha_weight: float = 0.0     # alpha towards 0 on sky-mask pixels
    fill_weight: float = 0.0          # alpha towards 1 on the other pixels (cameras with a sky mask)
    use_masks: bool = False           # no gradient where the keep mask is 0
    learn_sky: bool = False           # direction-dependent background (equirect texture)
    sky_width: int = 512
    sky_height: int = 128
    sky_lr: float = 0.01
    exposure: bool = False            # per-image affine colour compensation
    exposure_lr: float = 5e-3
    exposure_reg: float = 1e-2
    max_scale_ratio: float = 0.0      # cap largest / middle scale (<= 1: off)

    def active(self) -> bool:
        """Model::priorsActive (model.cpp:686-691)."""
        return (self.depth_weight > 0 or self.sky_alpha_weight > 0 or self.fill_weight > 0 or self.use_masks or
                self.learn_sky or self.exposure or self.max_scale_ratio > 1.0)


def depth_weight_at(o: PriorOptions, step: int, max_steps: int) -> float:
    """Model::depthWeightAt (model.cpp:693-701) in float32."""
    w0, w1 = F(o.depth_weight), F(o.depth_weight_final)
    if w0 <= 0:
        return 0.0
    if w1 < 0:
        return float(w0)
    t = F(min(max(F(step) / F(max(1, max_steps)), F(0)), F(1)))
    if w1 == 0:
        return float(w0 * (F(1) - t))
    return float(expf(logf(w0) * (F(1) - t) + logf(w1) * t))


def sky_frame(cameras) -> np.ndarray:
    """computeSkyFrame (model.cpp:704-737): zenith = the mean camera up, azimuth 0 = the
    mean heading made orthogonal to it, azimuth 90 = up x heading (double, stored float)."""
    up, fwd = np.zeros(3), np.zeros(3)
    for c in cameras:
        m = np.asarray(c["cam_to_world"], np.float32).reshape(4, 4).astype(np.float64)
        up += m[:3, 1]
        fwd -= m[:3, 2]
    n = np.linalg.norm(up)
    up = up / n if n >= 1e-12 else np.array([0.0, 1.0, 0.0])
    fwd = fwd - np.dot(fwd, up) * up
    n = np.linalg.norm(fwd)
    if n < 1e-12:
        e = np.zeros(3)
        e[int(np.argmin(np.abs(up)))] = 1.0
        fwd = e - np.dot(e, up) * up
        fwd /= np.linalg.norm(fwd)
    else:
        fwd /= n
    e2 = np.array([up[1] * fwd[2] - up[2] * fwd[1], up[2] * fwd[0] - up[0] * fwd[2], up[0] * fwd[1] - up[1] * fwd[0]])
    return np.concatenate([up, fwd, e2]).astype(np.float32)


def init_sky_texture(cameras, images, camera_priors, frame, W: int, H: int) -> np.ndarray:
    """initSkyTexture (model.cpp:739-812): the sky-mask pixels of the training images
    (every stride-th, stride = max(1, max(w, h) / 256)) splatted to the texel their ray
    hits; texels no image saw take their row's mean, else the nearest row with data, else
    the mean sky colour ((0.6, 0.7, 0.85) when no sky was seen). (H, W, 3) float32."""
    fr = np.asarray(frame, np.float32)
    tot, cnt = np.zeros((H * W, 3)), np.zeros(H * W)
    for cam, img, pr in zip(cameras, images, camera_priors):
        if not pr or not pr.get("has_sky") or pr.get("aux") is None:
            continue
        aux = np.asarray(pr["aux"], np.uint8)
        ph, pw = aux.shape[:2]
        ih, iw = img.shape[:2]
        stride = max(1, max(iw, ih) // 256)
        ys, xs = np.meshgrid(np.arange(0, ih, stride), np.arange(0, iw, stride), indexing="ij")
        qy = np.minimum(ph - 1, ((ys + F(0.5)).astype(np.float32) * F(ph) / F(ih)).astype(np.int64))
        qx = np.minimum(pw - 1, ((xs + F(0.5)).astype(np.float32) * F(pw) / F(iw)).astype(np.int64))
        sky = aux[qy, qx, 1] >= 128
        ys, xs = ys[sky], xs[sky]
        m = np.asarray(cam["cam_to_world"], np.float32).reshape(4, 4)
        dc = np.stack([(xs + F(0.5) - F(cam["cx"])) / F(cam["fx"]), (ys + F(0.5) - F(cam["cy"])) / F(cam["fy"]),
                       np.ones(xs.shape, np.float32)], 1).astype(np.float32)
        d = (dc[:, :1] * m[:3, 0] - dc[:, 1:2] * m[:3, 1] - dc[:, 2:3] * m[:3, 2]).astype(np.float32)
        n = np.sqrt((d * d).sum(1)).astype(np.float32)
        su = (d @ fr[:3]) / n
        el = np.arcsin(np.clip(su, -1, 1)).astype(np.float32)
        az = np.arctan2(d @ fr[6:9], d @ fr[3:6]).astype(np.float32)
        u = np.floor((az / F(2 * np.pi) + F(0.5)) * F(W)).astype(np.int64) % W
        v = np.clip(np.floor((F(0.5) - el / F(np.pi)) * F(H)).astype(np.int64), 0, H - 1)
        t = v * W + u
        np.add.at(tot, t, img[ys, xs].astype(np.float64))
        np.add.at(cnt, t, 1.0)
    mean = (tot.sum(0) / cnt.sum()).astype(np.float32) if cnt.sum() > 0 else np.array([0.6, 0.7, 0.85], np.float32)
    tot, cnt = tot.reshape(H, W, 3), cnt.reshape(H, W)
    rows = cnt.sum(1) > 0
    row_mean = np.where(rows[:, None], tot.sum(1) / np.maximum(cnt.sum(1), 1)[:, None], 0).astype(np.float32)
    out = np.zeros((H, W, 3), np.float32)
    for v in range(H):
        fill = mean
        for dv in range(H):
            if v - dv >= 0 and rows[v - dv]:
                fill = row_mean[v - dv]
                break
            if v + dv < H and rows[v + dv]:
                fill = row_mean[v + dv]
                break
        seen = cnt[v] > 0
        out[v] = np.where(seen[:, None], tot[v] / np.maximum(cnt[v], 1)[:, None], fill)
    return out


def _powf(b: float, e: int) -> np.float32:
    """std::pow(float, float) as the reference's sky and exposure schedules call it."""
    from .camera import _libm
    if not hasattr(_libm.powf, "_set"):
        _libm.powf.argtypes, _libm.powf.restype, _libm.powf._set = [C.c_float, C.c_float], C.c_float, True
    return F(_libm.powf(float(F(b)), float(F(e))))


class Trainer:
    """Trains device-resident gaussians on (camera, target) pairs, one view per step.
    With `priors` (PriorOptions) and `camera_priors` (per camera None or a dict: depth
    (ph, pw) float32 or None, aux (ph, pw, 4) uint8 confidence/sky/keep, has_depth,
    has_sky, has_mask), steps run the prior-guided path as Model does (model.cpp:548-870)."""

    def __init__(self, ctx: Context, dg: DeviceGaussians, cameras, targets, config: TrainConfig,
                 priors: PriorOptions | None = None, camera_priors=None):
        self.ctx, self.dg, self.cams, self.targets, self.cfg = ctx, dg, list(cameras), list(targets), config
        self.step_count = 0          # current_step
        self.adam_step_count = 0
        self.lr = list(LR_INIT)
        self.rng = np.random.default_rng(config.seed)
        self.order, self.order_pos = [], 0
        self.bg = floats(config.bg_color)
        self.last_hw = (0, 0)
        self.stop_split_at = config.iterations // 2
        self.densify_log = []
        self._struct_cache = {}
        self._views = {}
        self._pyramids = None
        self.stats_defined = False
        ctx.set_step_graph(config.step_graph)
        self.priors = priors if priors is not None and priors.active() else None
        if self.priors is not None:
            self._configure_priors(list(camera_priors or [None] * len(self.cams)))

    # ── prior-guided training (Model::configurePriors / trainPriorStep, model.cpp:548-870)
    def _configure_priors(self, camera_priors):
        from .prior import DevicePrior
        o = self.priors
        self.cam_priors = []
        for pr in camera_priors:
            if not pr or pr.get("aux") is None:
                self.cam_priors.append(None)
                continue
            aux = np.ascontiguousarray(pr["aux"], np.uint8)
            depth = pr.get("depth")
            depth = np.ascontiguousarray(depth if depth is not None else np.zeros(aux.shape[:2]), np.float32)
            self.cam_priors.append(dict(depth=DeviceArray.from_numpy(self.ctx, depth),
                                        aux=DeviceArray.from_numpy(self.ctx, aux), w=aux.shape[1], h=aux.shape[0],
                                        has_depth=bool(pr.get("has_depth", depth.any())),
                                        has_sky=bool(pr.get("has_sky", False)), has_mask=bool(pr.get("has_mask", False))))
        arrays = {"loss_terms": np.zeros(4, np.float32)}
        self.sky_frame = None
        if o.learn_sky:
            if o.sky_width < 4 or o.sky_height < 2:
                raise ValueError("sky texture must be at least 4x2")
            self.sky_frame = sky_frame(self.cams)
            images = [t.numpy().reshape(c["H"], c["W"], 3) for c, t in zip(self.cams, self.targets)]
            tex = init_sky_texture(self.cams, images, camera_priors, self.sky_frame, o.sky_width, o.sky_height)
            z = np.zeros_like(tex)
            arrays.update(sky_tex=tex, sky_grad=z, sky_m1=z, sky_m2=z)
        if o.exposure:
            n = len(self.cams)
            ident = np.zeros((n, 12), np.float32)
            ident[:, [0, 4, 8]] = 1.0
            arrays.update(expo_params=ident, expo_m1=np.zeros((n, 12), np.float32),
                          expo_m2=np.zeros((n, 12), np.float32), expo_grad=np.zeros(12, np.float32))
        self.prior_state = DevicePrior(self.ctx, arrays)
        self.sky_steps = 0
        self.expo_steps = [0] * len(self.cams)

    def _render_prior(self, force_aux: bool) -> Prior:
        """renderPriorStep (model.cpp:549-561): the learned sky, if any."""
        o, st = self.priors, self.prior_state
        p = Prior()
        p.aux = 1 if force_aux else 0
        if o.learn_sky:
            p.aux, p.sky = 1, 1
            p.sky_tex = st.arrays["sky_tex"].ptr
            p.sky_w, p.sky_h = o.sky_width, o.sky_height
            p.sky_frame[:] = [float(x) for x in self.sky_frame]
        return p

    def _train_prior(self, i: int, step: int) -> Prior:
        """trainPriorStep (model.cpp:574-623) for camera i: advances the sky and exposure
        step counts (once per training step)."""
        o, st = self.priors, self.prior_state
        p = self._render_prior(False)
        cp = self.cam_priors[i]
        depth_w = depth_weight_at(o, step, self.cfg.iterations) if cp and cp["has_depth"] else 0.0
        sky_terms = bool(cp and cp["has_sky"] and (o.sky_alpha_weight > 0 or o.fill_weight > 0))
        masks = bool(cp and cp["has_mask"] and o.use_masks)
        p.aux = 1 if (p.aux or depth_w > 0 or sky_terms or masks) else 0
        b1, b2 = F(0.9), F(0.999)
        if p.sky:
            self.sky_steps += 1
            t = self.sky_steps
            p.sky_grad, p.sky_exp_avg, p.sky_exp_avg_sq = (st.arrays[k].ptr for k in ("sky_grad", "sky_m1", "sky_m2"))
            p.sky_step_size = float(F(o.sky_lr) / (F(1) - _powf(b1, t)))
            p.sky_bc2_sqrt = float(np.sqrt(F(1) - _powf(b2, t)))
        if cp:
            p.prior_depth, p.prior_aux = cp["depth"].ptr, cp["aux"].ptr
            p.prior_w, p.prior_h = cp["w"], cp["h"]
            p.has_depth, p.has_sky_mask = int(cp["has_depth"]), int(cp["has_sky"])
        p.depth_weight, p.sky_weight, p.fill_weight = depth_w, o.sky_alpha_weight, o.fill_weight
        p.huber_delta, p.min_alpha, p.mask_photometric = o.depth_huber_delta, o.depth_min_alpha, int(o.use_masks)
        p.loss_terms = st.arrays["loss_terms"].ptr
        if o.exposure:
            self.expo_steps[i] += 1
            t = self.expo_steps[i]
            p.exposure, p.cam_index, p.num_cams = 1, i, len(self.cams)
            p.expo_params, p.expo_exp_avg, p.expo_exp_avg_sq, p.expo_grad = (
                st.arrays[k].ptr for k in ("expo_params", "expo_m1", "expo_m2", "expo_grad"))
            p.expo_step_size = float(F(o.exposure_lr) / (F(1) - _powf(b1, t)))
            p.expo_bc2_sqrt = float(np.sqrt(F(1) - _powf(b2, t)))
            p.expo_reg = o.exposure_reg
        if o.max_scale_ratio > 1.0:
            p.log_max_scale_ratio = float(logf(o.max_scale_ratio))
        return p

    def render(self, i: int, depth: bool = False):
        """Model::render / renderDepth (model.cpp:563-572, 866-881): camera i's image with
        the learned sky behind the gaussians; with depth, (expected depth, alpha) instead."""
        from .backend import render as plain_render
        cam = self._view(i, 1)
        dtu = min(self.step_count // max(self.cfg.sh_degree_interval, 1), self.cfg.sh_degree, self.dg.degree)
        g, c = self.dg.struct(), self._camera(i, dtu, 1)
        H, W = cam["H"], cam["W"]
        if self.priors is None and not depth:
            ptr = plain_render(self.ctx, self.dg, dict(cam, dtu=dtu), self.bg)
            return self.ctx.download(ptr, (H, W, 3), np.float32)
        p = self._render_prior(depth) if self.priors is not None else Prior()
        p.aux = 1 if (p.aux or depth) else 0
        out = C.c_void_p()
        check(lib().msplat_hip_render_prior(self.ctx.ptr, C.byref(g), C.byref(c), self.bg, C.byref(p), C.byref(out)))
        if not depth:
            return self.ctx.download(out.value, (H, W, 3), np.float32)
        from .prior import aux_outputs, finalize_depth
        d, t = aux_outputs(self.ctx, H, W)
        return finalize_depth(self.ctx, d, t, 1.0)

    def prior_losses(self) -> dict:
        """Model::lastPriorLosses (model.cpp:883-889): the last step's depth, sky and fill terms."""
        if self.priors is None:
            return {"depth": 0.0, "sky": 0.0, "fill": 0.0}
        t = self.prior_state.numpy("loss_terms")
        return {"depth": float(t[0]), "sky": float(t[1]), "fill": float(t[2])}

    def next_camera(self) -> int:
        if self.order_pos >= len(self.order):
            self.order = list(self.rng.permutation(len(self.cams)))
            self.order_pos = 0
        i = self.order[self.order_pos]
        self.order_pos += 1
        return int(i)

    def _camera(self, i: int, dtu: int, sf: int = 1):
        """The view of camera i at downscale factor sf (Model::prepareCam, model.cpp:498-546:
        the pyramid level's size and intrinsics scaled by the per-axis size ratio)."""
        key = (i, dtu, sf)
        c = self._struct_cache.get(key)
        if c is None:
            cam = self._view(i, sf)
            c = self._struct_cache[key] = make_camera_struct(dict(cam, dtu=dtu), self.dg.degree)
        return c

    def _view(self, i: int, sf: int) -> dict:
        cam = self.cams[i]
        if sf <= 1:
            return cam
        v = self._views.get((i, sf))
        if v is None:
            from .camera import make_camera
            v = self._views[(i, sf)] = make_camera(cam["cam_to_world"], cam["W"], cam["H"], cam["fx"], cam["fy"],
                                                   cam["cx"], cam["cy"], downscale=sf)
        return v

    def _target(self, i: int, sf: int):
        """Camera i's target at downscale sf: the pyramid level (Camera::getGPUImage)."""
        if sf <= 1:
            return self.targets[i]
        if self._pyramids is None:
            self._pyramids = {}
        pyr = self._pyramids.get(i)
        if pyr is None:
            from .images import Pyramid
            pyr = self._pyramids[i] = Pyramid(self.ctx, self.targets[i])
        return pyr(sf)

    def downscale_factor(self, step: int) -> int:
        """Model::getDownscaleFactor (model.cpp:195-197, image_sizing.cpp:129-133)."""
        from .camera import schedule_downscale
        return schedule_downscale(step, self.cfg.num_downscales, self.cfg.resolution_schedule)

    def step(self, read_loss: bool = False):
        """One iteration: train step, LR schedule, densification. Returns the loss when
        read_loss (a sync), else None. The step's radii are not copied out: the fused
        densification statistics already use them (they stay in ctx.frame().radii)."""
        self.step_count += 1
        step = self.step_count
        i = self.next_camera()
        sf = self.downscale_factor(step)
        cam = self._view(i, sf)
        dtu = min(step // max(self.cfg.sh_degree_interval, 1), self.cfg.sh_degree, self.dg.degree)
        c = self._camera(i, dtu, sf)
        target = self._target(i, sf)
        self.last_hw = (cam["H"], cam["W"])
        self.adam_step_count += 1
        steps, bc2s = adam_coeffs(self.lr, self.adam_step_count)
        g, a, s = self.dg.struct(), self.dg.adam(steps, bc2s), self.dg.stats()
        H, W = cam["H"], cam["W"]
        inv_n = float(np.float32(1.0) / np.float32(H * W * 3))
        inv_max = float(np.float32(1.0) / np.float32(max(H, W)))
        if not self.stats_defined:
            self.dg.reset_stats()
            self.stats_defined = True
        loss = C.c_float(0.0)
        if self.priors is not None:
            p = self._train_prior(i, step)
            check(lib().msplat_hip_train_step_prior(self.ctx.ptr, C.byref(g), C.byref(c), self.bg,
                                                    C.c_void_p(target.ptr), self.cfg.ssim_weight, inv_n, C.byref(a),
                                                    C.byref(s), inv_max, None, C.byref(loss) if read_loss else None,
                                                    C.byref(p)))
        else:
            check(lib().msplat_hip_train_step(self.ctx.ptr, C.byref(g), C.byref(c), self.bg,
                                              C.c_void_p(target.ptr), self.cfg.ssim_weight, inv_n, C.byref(a),
                                              C.byref(s), inv_max, None,
                                              C.byref(loss) if read_loss else None))
        self.schedulers_step(step)
        self.after_train(step)
        return loss.value if read_loss else None

    def save_checkpoint(self, path):
        """Model::saveCheckpoint (model.cpp:311-370): the step, the Adam step count and
        learning rates, parameters and both moments."""
        from .scene_io import save_checkpoint
        save_checkpoint(self.ctx, path, self.dg, self.step_count, self.adam_step_count, self.lr, LR_INIT[0],
                        MEANS_LR_FINAL)

    @classmethod
    def from_checkpoint(cls, ctx: Context, path, cameras, targets, config: "TrainConfig") -> "Trainer":
        """A trainer resumed from a checkpoint (Model::loadCheckpoint, model.cpp:372-496):
        parameters and moments in fresh 4N-capacity buffers, the step, Adam step count and
        learning rates restored; the densification statistics restart from zero."""
        from .scene_io import load_checkpoint
        dg, info = load_checkpoint(ctx, path)
        tr = cls(ctx, dg, cameras, targets, config)
        tr.step_count, tr.adam_step_count, tr.lr = info["step"], info["adam_step_count"], list(info["adam_lr"])
        return tr

    def schedulers_step(self, step: int):
        self.lr[0] = means_lr(step, self.cfg.iterations)

    def after_train(self, step: int):
        cfg = self.cfg
        if not cfg.densify:
            return
        f = refine_flags(step, cfg, len(self.cams), self.stop_split_at)
        if not f:
            return
        if f & REFINE_DENSIFY:
            before = self.dg.n
            self.dg.ensure_capacity(3 * self.dg.n)
            half_max_dim = float(F(0.5) * F(max(self.last_hw)))
            new = self.dg.densify(cfg.densify_grad_thresh, cfg.densify_size_thresh, cfg.split_screen_size,
                                  1 if f & REFINE_SCREEN else 0, cfg.cull_alpha_thresh, cfg.cull_scale_thresh,
                                  cfg.cull_screen_size, 1 if f & REFINE_HUGE else 0, half_max_dim,
                                  (step * 2654435761) % (1 << 32))
            self.densify_log.append((step, before, new))
        if f & REFINE_RESET:
            self.dg.opacity_reset(RESET_LOGIT)
        self.stats_defined = False  # statistics restart from zero (model.cpp:242-244)
