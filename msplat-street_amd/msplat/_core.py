"""The names, signatures and defaults of msplat's Python extension module `_core`
(python/bindings.cpp:431-778), on the HIP backend: TrainingConfig, TrainingStats, Dataset,
GaussianTrainer, the `_gpu_*` primitive hooks, `sync` and `cleanup`.

Two parts of the reference's surface are not here:
- Loading datasets from disk (COLMAP / Nerfstudio / Polycam folders, image decoding,
  `preflight`, sizing reports) is dataset I/O. A Dataset is built from arrays with
  `Dataset.from_arrays` instead: poses, intrinsics, decoded images and the initial
  point cloud. `Dataset(path, ...)` raises NotImplementedError.
- The prior-guided options (depth/sky/fill losses, masks, learned sky, exposure, scale
  cap) and the methods that read their state (`render_depth`, `prior_losses`,
  `export_sky`) raise NotImplementedError when used.
"""
from __future__ import annotations

import dataclasses
import time

import numpy as np

from ._abi import Context, DeviceArray, Prior, check, lib, make_camera_struct

_ctx: Context | None = None


def _context() -> Context:
    global _ctx
    if _ctx is None:
        _ctx = Context(0)
    return _ctx


def sync():
    """Synchronize GPU (wait for all commands to complete)."""
    if _ctx is not None:
        _ctx.sync()


def cleanup():
    """Release all cached GPU resources."""
    global _ctx
    if _ctx is not None:
        _ctx.close()
        _ctx = None


# ── TrainingConfig / TrainingStats (bindings.cpp:434-580) ──────────────────────

_PRIOR_FIELDS = ("depth_weight", "sky_alpha_weight", "fill_weight", "use_masks", "learn_sky",
                 "exposure_compensation")


@dataclasses.dataclass
class TrainingConfig:
    iterations: int = 30000
    sh_degree: int = 3
    sh_degree_interval: int = 1000
    ssim_weight: float = 0.2
    num_downscales: int = 2
    resolution_schedule: int = 3000
    refine_every: int = 100
    warmup_length: int = 500
    reset_alpha_every: int = 30
    densify_grad_thresh: float = 0.0002
    densify_size_thresh: float = 0.01
    stop_screen_size_at: int = 4000
    split_screen_size: float = 0.05
    keep_crs: bool = False
    downscale_factor: float = 1.0
    output: str = "splat.ply"
    save_every: int = -1
    bg_color: list = dataclasses.field(default_factory=lambda: [0.6130, 0.0101, 0.3984])
    depth_weight: float = 0.0
    depth_weight_final: float = -1.0
    depth_huber_delta: float = 0.05
    depth_min_alpha: float = 0.25
    sky_alpha_weight: float = 0.0
    fill_weight: float = 0.0
    use_masks: bool = False
    learn_sky: bool = False
    sky_width: int = 512
    sky_height: int = 128
    sky_lr: float = 0.01
    exposure_compensation: bool = False
    exposure_lr: float = 5e-3
    exposure_reg: float = 1e-2
    max_scale_ratio: float = 0.0

    def __post_init__(self):
        if len(self.bg_color) != 3:
            raise ValueError("bg_color must have exactly 3 elements [R, G, B]")
        if self.learn_sky and (self.sky_width < 4 or self.sky_height < 2):
            raise ValueError("sky texture must be at least 4x2")

    def priors_enabled(self) -> bool:
        return any(bool(getattr(self, f)) for f in _PRIOR_FIELDS) or self.max_scale_ratio > 1.0


@dataclasses.dataclass(frozen=True)
class TrainingStats:
    """Per-step training statistics returned by GaussianTrainer.step()."""
    iteration: int
    splat_count: int
    ms_per_step: float

    def __repr__(self):
        return f"TrainingStats(iteration={self.iteration}, splats={self.splat_count}, ms={self.ms_per_step:.6f})"


# ── Dataset ───────────────────────────────────────────────────────────────────

class Dataset:
    """Training (and held-out test) cameras with their target images and the initial
    point cloud. Build it with Dataset.from_arrays."""

    def __init__(self, path: str, downscale_factor: float = 1.0, eval_mode: bool = False, test_every: int = 8,
                 prior_dir: str = ""):
        raise NotImplementedError("loading datasets from disk (COLMAP, Nerfstudio, Polycam) is not part of this "
                                  "backend: build the Dataset with Dataset.from_arrays")

    @classmethod
    def from_arrays(cls, cam_to_world, intrinsics, images, points, colors, eval_mode: bool = False,
                    test_every: int = 8, priors=None) -> "Dataset":
        """cam_to_world: (n, 4, 4) OpenGL poses; intrinsics: (n, 6) rows (W, H, fx, fy, cx, cy);
        images: n decoded images, (H, W, 4) uint8 RGBA or (H, W, 3) float32, resampled to
        (W, H) on the GPU; points (m, 3) float32 and colors (m, 3) uint8: the initial cloud.
        With eval_mode every test_every-th camera is held out (as the loaders split).
        priors: per camera None or the geometric priors the reference loads from prior_dir
        (Camera::loadPriors): a dict with depth (ph, pw) float32 (0 = invalid) or None, aux
        (ph, pw, 4) uint8 (confidence, sky, keep, -), has_depth, has_sky, has_mask."""
        from .camera import make_camera
        from .images import load_target, resize_area
        ds = cls.__new__(cls)
        ctx = _context()
        cams, tgts = [], []
        for pose, intr, img in zip(cam_to_world, intrinsics, images):
            W, H, fx, fy, cx, cy = (float(v) for v in intr)
            W, H = int(W), int(H)
            cams.append(make_camera(np.asarray(pose, np.float32), W, H, fx, fy, cx, cy))
            img = np.asarray(img)
            if img.dtype == np.uint8:
                tgts.append(load_target(ctx, img, W, H))
            else:
                src = DeviceArray.from_numpy(ctx, np.ascontiguousarray(img, np.float32))
                tgts.append(src if img.shape[:2] == (H, W) else resize_area(ctx, src, W, H))
        test = [i for i in range(len(cams)) if eval_mode and i % test_every == 0]
        train = [i for i in range(len(cams)) if i not in test]
        ds._train = [(cams[i], tgts[i]) for i in train]
        ds._test = [(cams[i], tgts[i]) for i in test]
        pri = list(priors) if priors is not None else [None] * len(cams)
        ds._train_priors = [pri[i] for i in train]
        ds.points = np.ascontiguousarray(points, np.float32)
        ds.colors = np.ascontiguousarray(colors, np.uint8)
        return ds

    @property
    def num_train(self) -> int:
        return len(self._train)

    @property
    def num_test(self) -> int:
        return len(self._test)

    def camera_pose(self, index: int) -> np.ndarray:
        """Camera-to-world pose (4x4 row-major, OpenGL convention)."""
        return np.array(self._train[index][0]["cam_to_world"], np.float32).reshape(4, 4)

    def image(self, index: int, use_test: bool = False) -> np.ndarray:
        """The target image as loaded: numpy (H, W, 3) float32."""
        return (self._test if use_test else self._train)[index][1].numpy()


def _write_png(path: str, rgb: np.ndarray):
    """An 8-bit RGB PNG (no filtering, zlib-deflated scanlines)."""
    import struct
    import zlib
    h, w = rgb.shape[:2]
    raw = b"".join(b"\x00" + rgb[y].tobytes() for y in range(h))

    def chunk(tag, data):
        return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)
    with open(path, "wb") as f:
        f.write(b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 2, 0, 0, 0)) +
                chunk(b"IDAT", zlib.compress(raw, 6)) + chunk(b"IEND", b""))


# ── GaussianTrainer (bindings.cpp:598-653) ───────────────────────────────────

class GaussianTrainer:
    """3D Gaussian Splatting trainer. All computation runs on the GPU."""

    def __init__(self, dataset: Dataset, config: TrainingConfig):
        from .backend import DeviceGaussians
        from .trainer import PriorOptions, TrainConfig, Trainer
        self.dataset, self.config = dataset, config
        self.ctx = _context()
        dg = DeviceGaussians.from_points(self.ctx, dataset.points, dataset.colors, degree=config.sh_degree,
                                         capacity=4 * len(dataset.points))
        tc = TrainConfig(iterations=config.iterations, sh_degree=config.sh_degree,
                         sh_degree_interval=config.sh_degree_interval, ssim_weight=config.ssim_weight,
                         num_downscales=config.num_downscales, resolution_schedule=config.resolution_schedule,
                         refine_every=config.refine_every, warmup_length=config.warmup_length,
                         reset_alpha_every=config.reset_alpha_every, densify_grad_thresh=config.densify_grad_thresh,
                         densify_size_thresh=config.densify_size_thresh,
                         stop_screen_size_at=config.stop_screen_size_at, split_screen_size=config.split_screen_size,
                         bg_color=tuple(config.bg_color))
        po = None
        if config.priors_enabled():  # Model::configurePriors (model.cpp:814-864)
            po = PriorOptions(depth_weight=config.depth_weight, depth_weight_final=config.depth_weight_final,
                              depth_huber_delta=config.depth_huber_delta, depth_min_alpha=config.depth_min_alpha,
                              sky_alpha_weight=config.sky_alpha_weight, fill_weight=config.fill_weight,
                              use_masks=config.use_masks, learn_sky=config.learn_sky, sky_width=config.sky_width,
                              sky_height=config.sky_height, sky_lr=config.sky_lr,
                              exposure=config.exposure_compensation, exposure_lr=config.exposure_lr,
                              exposure_reg=config.exposure_reg, max_scale_ratio=config.max_scale_ratio)
        self._tr = Trainer(self.ctx, dg, [c for c, _ in dataset._train], [t for _, t in dataset._train], tc,
                           priors=po, camera_priors=getattr(dataset, "_train_priors", None))

    # training
    @property
    def current_step(self) -> int:
        return self._tr.step_count

    @property
    def iteration(self) -> int:
        return self._tr.step_count

    @property
    def splat_count(self) -> int:
        return self._tr.dg.n

    def step(self) -> TrainingStats:
        """Run a single training iteration. Returns TrainingStats."""
        t0 = time.perf_counter()
        self._tr.step()
        ms = (time.perf_counter() - t0) * 1e3
        if self.config.save_every > 0 and self._tr.step_count % self.config.save_every == 0:
            self.export_ply(self.config.output)
        return TrainingStats(self._tr.step_count, self._tr.dg.n, ms)

    def train(self, callback, callback_every: int = 100):
        """Run training to completion, calling callback(stats) every callback_every steps."""
        while self._tr.step_count < self.config.iterations:
            stats = self.step()
            if callback is not None and callback_every > 0 and stats.iteration % callback_every == 0:
                callback(stats)

    def resolution(self, cam_idx: int = 0, step: int = -1) -> tuple:
        """(width, height, factor) a training camera renders at at a step (default: the current one)."""
        from .camera import level_size
        f = self._tr.downscale_factor(self._tr.step_count if step < 0 else step)
        c = self.dataset._train[cam_idx][0]
        return level_size(c["W"], f), level_size(c["H"], f), f

    # rendering
    def _render(self, cam: dict) -> int:
        """Model::render (model.cpp:563-572): with a learned sky, the sky behind the gaussians."""
        import ctypes as C
        from .backend import render
        tr = self._tr
        dtu = min(tr.step_count // max(self.config.sh_degree_interval, 1), self.config.sh_degree)
        if tr.priors is not None and tr.priors.learn_sky:
            p = tr._render_prior(False)
            g, c = tr.dg.struct(), make_camera_struct(dict(cam, dtu=dtu), tr.dg.degree)
            out = C.c_void_p()
            check(lib().msplat_hip_render_prior(self.ctx.ptr, C.byref(g), C.byref(c), tr.bg, C.byref(p),
                                                C.byref(out)))
            return out.value
        return render(self.ctx, tr.dg, dict(cam, dtu=dtu), tuple(self.config.bg_color))

    def _cam(self, cam_idx: int, use_test: bool) -> dict:
        return (self.dataset._test if use_test else self.dataset._train)[cam_idx][0]

    def _pose_cam(self, cam_to_world, ref_cam_idx: int) -> dict:
        from .camera import make_camera
        r = self.dataset._train[ref_cam_idx][0]
        return make_camera(np.asarray(cam_to_world, np.float32).reshape(4, 4), r["W"], r["H"], r["fx"], r["fy"],
                           r["cx"], r["cy"])

    def _image(self, cam: dict) -> np.ndarray:
        ptr = self._render(cam)
        return self.ctx.download(ptr, (cam["H"], cam["W"], 3), np.float32)

    def _rgba8(self, cam: dict) -> np.ndarray:
        import ctypes as C
        ptr = self._render(cam)
        n = cam["W"] * cam["H"]
        out = DeviceArray(self.ctx, (n, 4), np.uint8, zero=False)
        check(lib().msplat_hip_pack_rgba8(self.ctx.ptr, C.c_void_p(ptr), n, C.c_void_p(out.ptr)))
        img = out.numpy().reshape(cam["H"], cam["W"], 4)
        out.free()
        return img

    def render(self, cam_idx: int, use_test: bool = False) -> np.ndarray:
        """Render a camera view: numpy (H, W, 3) float32, RGB [0, 1]."""
        return self._image(self._cam(cam_idx, use_test))

    def render_from_pose(self, cam_to_world, ref_cam_idx: int = 0) -> np.ndarray:
        """Render from a camera-to-world pose (4x4 row-major, OpenGL), intrinsics of ref_cam_idx."""
        return self._image(self._pose_cam(cam_to_world, ref_cam_idx))

    def render_rgba8(self, cam_idx: int, use_test: bool = False) -> np.ndarray:
        """Render a camera view for display: numpy (H, W, 4) uint8 RGBA, packed on the GPU."""
        return self._rgba8(self._cam(cam_idx, use_test))

    def render_from_pose_rgba8(self, cam_to_world, ref_cam_idx: int = 0) -> np.ndarray:
        return self._rgba8(self._pose_cam(cam_to_world, ref_cam_idx))

    def evaluate(self) -> dict:
        """Mean PSNR, SSIM (clamp-to-edge) and L1 over the held-out test cameras."""
        import ctypes as C
        if not self.dataset._test:
            raise RuntimeError("no test cameras: load the dataset with eval_mode=True")
        acc = np.zeros(3)
        out = (C.c_double * 3)()
        for cam, tgt in self.dataset._test:
            ptr = self._render(cam)
            check(lib().msplat_hip_image_metrics(self.ctx.ptr, C.c_void_p(ptr), C.c_void_p(tgt.ptr), cam["H"],
                                                 cam["W"], out))
            acc += np.array(out[:])
        acc /= len(self.dataset._test)
        return {"psnr": float(acc[0]), "ssim": float(acc[1]), "l1": float(acc[2])}

    # files
    def export_ply(self, path: str):
        from .scene_io import save_ply
        save_ply(self.ctx, self._tr.dg, path, step=self._tr.step_count)

    def export_splat(self, path: str):
        from .scene_io import save_splat
        save_splat(self.ctx, self._tr.dg, path)

    def save_checkpoint(self, path: str):
        self._tr.save_checkpoint(path)

    def load_checkpoint(self, path: str) -> int:
        """Load a training checkpoint and resume from the saved iteration."""
        from .trainer import Trainer
        self._tr = Trainer.from_checkpoint(self.ctx, path, self._tr.cams, self._tr.targets, self._tr.cfg)
        return self._tr.step_count

    # the prior-guided path
    def render_depth(self, cam_idx: int, use_test: bool = False):
        """Expected depth (dataset units, 0 where empty) and accumulated alpha: two numpy
        (H, W) float32 arrays (Model::renderDepth, model.cpp:866-881)."""
        import ctypes as C
        from .prior import aux_outputs, finalize_depth
        cam = self._cam(cam_idx, use_test)
        tr = self._tr
        dtu = min(tr.step_count // max(self.config.sh_degree_interval, 1), self.config.sh_degree)
        p = tr._render_prior(True) if tr.priors is not None else Prior()
        p.aux = 1
        g, c = tr.dg.struct(), make_camera_struct(dict(cam, dtu=dtu), tr.dg.degree)
        out = C.c_void_p()
        check(lib().msplat_hip_render_prior(self.ctx.ptr, C.byref(g), C.byref(c), tr.bg, C.byref(p), C.byref(out)))
        d, t = aux_outputs(self.ctx, cam["H"], cam["W"])
        return finalize_depth(self.ctx, d, t, 1.0)

    def prior_losses(self) -> dict:
        """Mean prior losses of the last step: dict with depth, sky, fill. Syncs the GPU."""
        return self._tr.prior_losses()

    def export_sky(self, path: str):
        """Save the learned sky as an equirect PNG (requires learn_sky=True): 8-bit RGB,
        round(clamp(v * 255, 0, 255)) as the reference's imwriteRGB."""
        tr = self._tr
        if tr.priors is None or not tr.priors.learn_sky:
            raise RuntimeError("No learned sky (enable learn_sky)")
        tex = tr.prior_state.numpy("sky_tex")
        rgb = (np.clip(tex * np.float32(255.0), 0, 255) + np.float32(0.5)).astype(np.uint8)
        _write_png(path, rgb)


# ── GPU primitive hooks (bindings.cpp:655-738) ───────────────────────────────

def _gpu_radix_sort(keys, key_bits: int = 64):
    """Stable GPU radix sort of uint64 keys. Returns (sorted_keys, order)."""
    ctx = _context()
    keys = np.ascontiguousarray(keys, np.uint64)
    n = keys.size
    if n == 0:
        return keys.copy(), np.zeros(0, np.uint32)
    dk, dv = DeviceArray.from_numpy(ctx, keys), DeviceArray.from_numpy(ctx, np.arange(n, dtype=np.int32))
    check(lib().msplat_hip_radix_sort(ctx.ptr, dk.ptr, dv.ptr, n, int(key_bits)))
    return dk.numpy(), dv.numpy().view(np.uint32)


def _gpu_knn3_mean_dist(points) -> np.ndarray:
    """Exact mean distance to the 3 nearest other points, computed on the GPU."""
    pts = np.ascontiguousarray(points, np.float32)
    if pts.ndim != 2 or pts.shape[1] != 3:
        raise ValueError("points must be (n, 3)")
    if pts.shape[0] == 0:
        raise ValueError("points must not be empty")
    ctx = _context()
    dp, out = DeviceArray.from_numpy(ctx, pts), DeviceArray(ctx, (pts.shape[0],))
    check(lib().msplat_hip_knn3_mean_dist(ctx.ptr, dp.ptr, pts.shape[0], out.ptr))
    return out.numpy()


def _gpu_resize_area(img, out_w: int, out_h: int) -> np.ndarray:
    """Area resample on the GPU: (H, W, 4) uint8 or (H, W, 3) float32 -> (out_h, out_w, 3) float32."""
    from .images import resize_area
    img = np.ascontiguousarray(img)
    if out_w < 1 or out_h < 1:
        raise ValueError("output size must be positive")
    if not (img.dtype == np.uint8 and img.ndim == 3 and img.shape[2] == 4) and \
            not (img.dtype == np.float32 and img.ndim == 3 and img.shape[2] == 3):
        raise ValueError("img must be (H, W, 4) uint8 or (H, W, 3) float32")
    ctx = _context()
    return resize_area(ctx, DeviceArray.from_numpy(ctx, img), int(out_w), int(out_h)).numpy()


def undistort_roi(w: int, h: int, fx, fy, cx, cy, k1, k2, p1, p2, k3) -> tuple:
    """The alpha = 0 crop of an undistorted image (undistortROI, image_sizing.cpp:33-69), in
    float32: the undistorted centres of 200 samples along each border, inverted from the
    Brown-Conrady model by 20 fixed-point iterations; returns (x, y, width, height)."""
    F = np.float32
    fx, fy, cx, cy, k1, k2, p1, p2, k3 = (F(v) for v in (fx, fy, cx, cy, k1, k2, p1, p2, k3))

    def undistort(px, py):
        xd, yd = (F(px) + F(0.5) - cx) / fx, (F(py) + F(0.5) - cy) / fy
        xu, yu = xd, yd
        for _ in range(20):
            r2 = xu * xu + yu * yu
            r4 = r2 * r2
            r6 = r4 * r2
            radial = F(1) + k1 * r2 + k2 * r4 + k3 * r6
            dx = F(2) * p1 * xu * yu + p2 * (r2 + F(2) * xu * xu)
            dy = p1 * (r2 + F(2) * yu * yu) + F(2) * p2 * xu * yu
            xu, yu = (xd - dx) / radial, (yd - dy) / radial
        return xu * fx + cx - F(0.5), yu * fy + cy - F(0.5)

    top, bottom, left, right = F(-1e9), F(1e9), F(-1e9), F(1e9)
    for i in range(200):
        t = F(i) / F(199)
        top = max(top, undistort(t * F(w - 1), F(0))[1])
        bottom = min(bottom, undistort(t * F(w - 1), F(h - 1))[1])
        left = max(left, undistort(F(0), t * F(h - 1))[0])
        right = min(right, undistort(F(w - 1), t * F(h - 1))[0])
    x, y = max(0, int(np.ceil(left))), max(0, int(np.ceil(top)))
    cw, ch = min(w, int(np.floor(right))) - x, min(h, int(np.floor(bottom))) - y
    if cw <= 0 or ch <= 0:
        return 0, 0, w, h
    return x, y, cw, ch


def _gpu_undistort(img, intrinsics, distortion) -> tuple:
    """Undistort on the GPU into the alpha = 0 crop. Returns (image, roi_x, roi_y)."""
    from .images import undistort
    if len(intrinsics) != 4 or len(distortion) != 5:
        raise ValueError("intrinsics = [fx, fy, cx, cy], distortion = [k1, k2, p1, p2, k3]")
    img = np.ascontiguousarray(img, np.float32)
    h, w = img.shape[:2]
    x, y, cw, ch = undistort_roi(w, h, *intrinsics, *distortion)
    ctx = _context()
    out = undistort(ctx, DeviceArray.from_numpy(ctx, img), intrinsics, distortion, x, y, cw, ch)
    return out.numpy(), x, y


def _gpu_image_metrics(rendered, gt) -> tuple:
    """GPU (psnr, ssim, l1) of two (H, W, 3) float32 images."""
    import ctypes as C
    a, b = np.ascontiguousarray(rendered, np.float32), np.ascontiguousarray(gt, np.float32)
    if a.shape != b.shape or a.ndim != 3 or a.shape[2] != 3:
        raise ValueError("images must have the same (H, W, 3) size")
    ctx = _context()
    da, db = DeviceArray.from_numpy(ctx, a), DeviceArray.from_numpy(ctx, b)
    out = (C.c_double * 3)()
    check(lib().msplat_hip_image_metrics(ctx.ptr, C.c_void_p(da.ptr), C.c_void_p(db.ptr), a.shape[0], a.shape[1],
                                         out))
    return tuple(out[:])


def _gpu_pack_rgba8(img) -> np.ndarray:
    """Pack (H, W, 3) float32 into (H, W, 4) uint8 RGBA on the GPU."""
    import ctypes as C
    a = np.ascontiguousarray(img, np.float32)
    ctx = _context()
    n = a.shape[0] * a.shape[1]
    da, out = DeviceArray.from_numpy(ctx, a), DeviceArray(ctx, (max(n, 1), 4), np.uint8, zero=False)
    check(lib().msplat_hip_pack_rgba8(ctx.ptr, C.c_void_p(da.ptr), n, C.c_void_p(out.ptr)))
    return out.numpy()[:n].reshape(a.shape[0], a.shape[1], 4)
