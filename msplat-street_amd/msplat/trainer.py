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

from ._abi import Context, DeviceArray, check, floats, lib, make_camera_struct
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


class Trainer:
    """Trains device-resident gaussians on (camera, target) pairs, one view per step."""

    def __init__(self, ctx: Context, dg: DeviceGaussians, cameras, targets, config: TrainConfig):
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
