"""ctypes binding of the HIP backend's C ABI (include/msplat_hip.h).

Loads msplat/libmsplat_hip.so built in-tree; there is no fallback: if the library or a
GPU is missing, every call raises. DeviceArray wraps one device allocation with numpy
upload/download.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmsplat_hip.so")

MSPLAT_OK = 0
FRAME_KEEP_GRADS = 1
FRAME_KEYS64 = 2
TUNE_BWD_PPT = 0
TUNE_FWD_BATCH = 1
TUNE_BWD_BATCH = 2
TUNE_CULL = 3
TUNE_PACK_KEYS = 4
TUNE_CHUNKS = 5
# stage timing: stage s is bit (1 << s) of a stage mask
STAGE_NAMES = ("proj_sh_fwd", "prefix_sort_pack", "rast_fwd", "loss_fwd_bwd", "rast_bwd", "proj_sh_bwd_adam")
fp = C.POINTER(C.c_float)


class MsplatError(RuntimeError):
    pass


class Camera(C.Structure):
    _fields_ = [("viewmat", C.c_float * 16), ("projmat", C.c_float * 16), ("cam_pos", C.c_float * 3),
                ("fx", C.c_float), ("fy", C.c_float), ("cx", C.c_float), ("cy", C.c_float),
                ("width", C.c_uint32), ("height", C.c_uint32), ("tiles_x", C.c_int32), ("tiles_y", C.c_int32),
                ("clip_thresh", C.c_float), ("glob_scale", C.c_float), ("degree", C.c_uint32),
                ("degrees_to_use", C.c_uint32)]


class Gaussians(C.Structure):
    _fields_ = [("num_points", C.c_int32), ("means", C.c_void_p), ("scales", C.c_void_p), ("quats", C.c_void_p),
                ("features_dc", C.c_void_p), ("features_rest", C.c_void_p), ("opacities", C.c_void_p)]


class Adam(C.Structure):
    _fields_ = [("exp_avg", C.c_void_p * 6), ("exp_avg_sq", C.c_void_p * 6), ("step_size", C.c_float * 6),
                ("bc2_sqrt", C.c_float * 6), ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float)]


class Prior(C.Structure):
    """msplat_hip_prior: PriorStep (core/metal/bindings.h:43-77) with device pointers."""
    _fields_ = [("aux", C.c_int32), ("sky", C.c_int32), ("sky_tex", C.c_void_p), ("sky_grad", C.c_void_p),
                ("sky_exp_avg", C.c_void_p), ("sky_exp_avg_sq", C.c_void_p), ("sky_w", C.c_int32),
                ("sky_h", C.c_int32), ("sky_frame", C.c_float * 9), ("sky_step_size", C.c_float),
                ("sky_bc2_sqrt", C.c_float), ("prior_depth", C.c_void_p), ("prior_aux", C.c_void_p),
                ("prior_w", C.c_int32), ("prior_h", C.c_int32), ("has_depth", C.c_int32),
                ("has_sky_mask", C.c_int32), ("depth_weight", C.c_float), ("sky_weight", C.c_float),
                ("fill_weight", C.c_float), ("huber_delta", C.c_float), ("min_alpha", C.c_float),
                ("mask_photometric", C.c_int32), ("loss_terms", C.c_void_p), ("exposure", C.c_int32),
                ("cam_index", C.c_int32), ("num_cams", C.c_int32), ("expo_params", C.c_void_p),
                ("expo_exp_avg", C.c_void_p), ("expo_exp_avg_sq", C.c_void_p), ("expo_grad", C.c_void_p),
                ("expo_step_size", C.c_float), ("expo_bc2_sqrt", C.c_float), ("expo_reg", C.c_float),
                ("log_max_scale_ratio", C.c_float)]


class GradStats(C.Structure):
    _fields_ = [("vis_counts", C.c_void_p), ("xys_grad_norm", C.c_void_p), ("max_2d_size", C.c_void_p)]


class Grads(C.Structure):
    _fields_ = [("v_means", C.c_void_p), ("v_scales", C.c_void_p), ("v_quats", C.c_void_p),
                ("v_features_dc", C.c_void_p), ("v_features_rest", C.c_void_p), ("v_opacities", C.c_void_p)]


class Views(C.Structure):
    _fields_ = [("num_views", C.c_int32), ("cameras", C.POINTER(Camera)), ("backgrounds", C.POINTER(C.c_float)),
                ("targets", C.POINTER(C.c_void_p))]


MAX_BATCH_VIEWS = 64


class Frame(C.Structure):
    _fields_ = [("num_points", C.c_int32), ("num_tiles", C.c_int32), ("num_isects", C.c_int64),
                ("splat_a", C.c_void_p), ("splat_b", C.c_void_p), ("splat_blue", C.c_void_p),
                ("depths", C.c_void_p), ("aabb", C.c_void_p), ("radii", C.c_void_p), ("tile_count", C.c_void_p),
                ("offsets", C.c_void_p), ("isect_keys", C.c_void_p), ("isect_gid", C.c_void_p),
                ("tile_start", C.c_void_p), ("image", C.c_void_p), ("final_T", C.c_void_p),
                ("v_image", C.c_void_p), ("final_idx", C.c_void_p), ("v_xy", C.c_void_p),
                ("v_conic", C.c_void_p), ("v_rgb", C.c_void_p), ("v_opacity", C.c_void_p),
                ("loss_sum", C.c_void_p), ("slots_written", C.c_void_p), ("num_visible", C.c_int64),
                ("depth_chunks", C.c_int32), ("isects_running", C.c_void_p), ("visible_running", C.c_void_p)]


class Checkpoint(C.Structure):
    _fields_ = [("version", C.c_int32), ("step", C.c_int32), ("sh_degree", C.c_int32), ("adam_step_count", C.c_int32),
                ("adam_lr", C.c_float * 6), ("means_lr_init", C.c_float), ("means_lr_final", C.c_float),
                ("gaussians", Gaussians), ("exp_avg", C.c_void_p * 6), ("exp_avg_sq", C.c_void_p * 6),
                ("has_sky", C.c_int32), ("has_exposure", C.c_int32)]


_lib = None


def lib():
    """The loaded library (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MsplatError(f"{LIB_PATH} is not built: run python msplat-street_amd/build.py")
        L = C.CDLL(LIB_PATH)
        vp, i32, u32, f32 = C.c_void_p, C.c_int32, C.c_uint32, C.c_float
        sig = {
            "msplat_hip_last_error": ([], C.c_char_p),
            "msplat_hip_ctx_create": ([C.c_int, vp, C.POINTER(vp)], C.c_int),
            "msplat_hip_ctx_destroy": ([vp], C.c_int),
            "msplat_hip_ctx_set_stream": ([vp, vp], C.c_int),
            "msplat_hip_ctx_stream": ([vp], vp),
            "msplat_hip_sync": ([vp], C.c_int),
            "msplat_hip_commit": ([vp], C.c_int),
            "msplat_hip_malloc": ([C.c_size_t, C.c_int, C.POINTER(vp)], C.c_int),
            "msplat_hip_free": ([vp], C.c_int),
            "msplat_hip_memcpy": ([vp, vp, vp, C.c_size_t, C.c_int], C.c_int),
            "msplat_hip_memset": ([vp, vp, C.c_int, C.c_size_t], C.c_int),
            "msplat_hip_render": ([vp, C.POINTER(Gaussians), C.POINTER(Camera), fp, C.POINTER(vp)], C.c_int),
            "msplat_hip_train_step": ([vp, C.POINTER(Gaussians), C.POINTER(Camera), fp, vp, f32, f32,
                                       C.POINTER(Adam), C.POINTER(GradStats), f32, vp, fp], C.c_int),
            "msplat_hip_last_loss": ([vp, fp], C.c_int),
            "msplat_hip_train_step_prior": ([vp, C.POINTER(Gaussians), C.POINTER(Camera), fp, vp, f32, f32,
                                             C.POINTER(Adam), C.POINTER(GradStats), f32, vp, fp, C.POINTER(Prior)],
                                            C.c_int),
            "msplat_hip_render_prior": ([vp, C.POINTER(Gaussians), C.POINTER(Camera), fp, C.POINTER(Prior),
                                         C.POINTER(vp)], C.c_int),
            "msplat_hip_render_aux_outputs": ([vp, C.POINTER(vp), C.POINTER(vp)], C.c_int),
            "msplat_hip_finalize_depth": ([vp, vp, vp, C.c_uint32, f32, vp, vp], C.c_int),
            "msplat_hip_render_backward": ([vp, C.POINTER(Gaussians), C.POINTER(Camera), fp, vp,
                                            C.POINTER(Grads)], C.c_int),
            "msplat_hip_frame_view": ([vp, C.POINTER(Frame)], C.c_int),
            "msplat_hip_set_flags": ([vp, C.c_int], C.c_int),
            "msplat_hip_stage_project": ([vp, C.POINTER(Gaussians), C.POINTER(Camera)], C.c_int),
            "msplat_hip_stage_bin": ([vp, C.POINTER(Camera)], C.c_int),
            "msplat_hip_stage_raster_forward": ([vp, C.POINTER(Camera), fp], C.c_int),
            "msplat_hip_stage_loss": ([vp, C.POINTER(Camera), vp, f32, f32, fp], C.c_int),
            "msplat_hip_device_memory": ([C.c_int, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)], C.c_int),
            "msplat_hip_resize_area": ([vp, vp, C.c_int, C.c_int, C.c_int, vp, C.c_int, C.c_int], C.c_int),
            "msplat_hip_undistort": ([vp, vp, C.c_int, C.c_int, fp, fp, C.c_int, C.c_int, vp, C.c_int, C.c_int], C.c_int),
            "msplat_hip_save_ply": ([vp, C.POINTER(Gaussians), C.c_int, C.c_char_p, C.c_int, f32, fp, C.c_int], C.c_int),
            "msplat_hip_save_splat": ([vp, C.POINTER(Gaussians), C.c_char_p, f32, fp, C.c_int], C.c_int),
            "msplat_hip_ply_info": ([C.c_char_p, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)], C.c_int),
            "msplat_hip_load_ply": ([vp, C.c_char_p, C.POINTER(Gaussians), C.c_int, f32, fp, C.c_int], C.c_int),
            "msplat_hip_save_checkpoint": ([vp, C.c_char_p, C.POINTER(Checkpoint)], C.c_int),
            "msplat_hip_checkpoint_info": ([C.c_char_p, C.POINTER(Checkpoint)], C.c_int),
            "msplat_hip_load_checkpoint": ([vp, C.c_char_p, C.POINTER(Checkpoint)], C.c_int),
            "msplat_hip_radix_sort": ([vp, vp, vp, u32, C.c_int], C.c_int),
            "msplat_hip_radix_sort_noalloc": ([vp, vp, vp, u32, C.c_int], C.c_int),
            "msplat_hip_densify": ([vp, i32, i32, f32, f32, f32, C.c_int, f32, f32, f32, C.c_int, vp, vp, vp, f32,
                                    C.c_void_p * 18, i32, u32, C.POINTER(i32)], C.c_int),
            "msplat_hip_opacity_reset": ([vp, vp, vp, vp, i32, f32], C.c_int),
            "msplat_hip_knn3_mean_dist": ([vp, vp, u32, vp], C.c_int),
            "msplat_hip_init_gaussians": ([vp, vp, vp, u32, u32, f32, vp, vp, vp, vp], C.c_int),
            "msplat_hip_image_metrics": ([vp, vp, vp, C.c_int, C.c_int, C.POINTER(C.c_double)], C.c_int),
            "msplat_hip_pack_rgba8": ([vp, vp, u32, vp], C.c_int),
            "msplat_hip_enable_stage_timing": ([vp, C.c_int], C.c_int),
            "msplat_hip_set_stage_mask": ([vp, C.c_uint32], C.c_int),
            "msplat_hip_drain_stage_times": ([vp, C.POINTER(C.c_double), C.POINTER(C.c_int64), C.c_int,
                                              C.POINTER(C.c_int)], C.c_int),
            "msplat_hip_stage_name": ([C.c_int], C.c_char_p),
            "msplat_hip_set_tuning": ([vp, C.c_int, C.c_int], C.c_int),
            "msplat_hip_get_tuning": ([vp, C.c_int, C.POINTER(C.c_int)], C.c_int),
            "msplat_hip_render_batch": ([vp, C.POINTER(Gaussians), C.POINTER(Views), C.POINTER(vp),
                                         C.POINTER(C.c_uint64)], C.c_int),
            "msplat_hip_train_step_batch": ([vp, C.POINTER(Gaussians), C.POINTER(Views), f32, C.POINTER(Adam),
                                             C.POINTER(GradStats), fp], C.c_int),
            "msplat_hip_render_backward_batch": ([vp, C.POINTER(Gaussians), C.POINTER(Views), vp,
                                                  C.POINTER(Grads)], C.c_int),
            "msplat_hip_set_batch_graph": ([vp, C.c_int], C.c_int),
            "msplat_hip_batch_graph_stats": ([vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)], C.c_int),
            "msplat_hip_set_step_graph": ([vp, C.c_int, C.c_uint64], C.c_int),
            "msplat_hip_step_graph_stats": ([vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)],
                                            C.c_int),
            "msplat_hip_batch_frame_info": ([vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)], C.c_int),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes, f.restype = args, res
        _lib = L
    return _lib


def check(rc: int):
    if rc != MSPLAT_OK:
        raise MsplatError(f"msplat HIP error {rc}: {lib().msplat_hip_last_error().decode()}")


class Context:
    """A HIP context: one stream plus the per-frame cache."""

    def __init__(self, device: int = 0, stream: int | None = None):
        self.ptr = C.c_void_p()
        check(lib().msplat_hip_ctx_create(device, C.c_void_p(stream) if stream else None, C.byref(self.ptr)))

    def close(self):
        if self.ptr:
            lib().msplat_hip_ctx_destroy(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        check(lib().msplat_hip_sync(self.ptr))

    def frame(self) -> Frame:
        f = Frame()
        check(lib().msplat_hip_frame_view(self.ptr, C.byref(f)))
        return f

    def set_flags(self, flags: int):
        check(lib().msplat_hip_set_flags(self.ptr, flags))

    def set_tuning(self, key: int, value: int):
        check(lib().msplat_hip_set_tuning(self.ptr, key, value))

    def get_tuning(self, key: int) -> int:
        v = C.c_int()
        check(lib().msplat_hip_get_tuning(self.ptr, key, C.byref(v)))
        return v.value

    def enable_stage_timing(self, on: bool = True):
        check(lib().msplat_hip_enable_stage_timing(self.ptr, 1 if on else 0))

    def set_stage_mask(self, stages=None):
        """Time only the named stages (None: every stage)."""
        mask = (1 << len(STAGE_NAMES)) - 1 if stages is None else sum(1 << STAGE_NAMES.index(s) for s in stages)
        check(lib().msplat_hip_set_stage_mask(self.ptr, mask))

    def drain_stage_times(self) -> dict:
        """{stage name: (total ms, launches)} since the last drain (syncs)."""
        ms = (C.c_double * 16)()
        calls = (C.c_int64 * 16)()
        n = C.c_int()
        check(lib().msplat_hip_drain_stage_times(self.ptr, ms, calls, 16, C.byref(n)))
        return {lib().msplat_hip_stage_name(i).decode(): (ms[i], calls[i]) for i in range(n.value)}

    def set_batch_graph(self, on: bool = True):
        check(lib().msplat_hip_set_batch_graph(self.ptr, 1 if on else 0))

    def set_step_graph(self, on: bool = True, initial_capacity: int = 0):
        """Replay msplat_hip_train_step from a captured hipGraph (msplat_hip_set_step_graph)."""
        check(lib().msplat_hip_set_step_graph(self.ptr, 1 if on else 0, int(initial_capacity)))

    def step_graph_stats(self) -> tuple:
        """(captures, replays, replays run again after an intersection overflow)"""
        c, r, x = C.c_int64(), C.c_int64(), C.c_int64()
        check(lib().msplat_hip_step_graph_stats(self.ptr, C.byref(c), C.byref(r), C.byref(x)))
        return c.value, r.value, x.value

    def batch_graph_stats(self) -> tuple:
        c, r = C.c_int64(), C.c_int64()
        check(lib().msplat_hip_batch_graph_stats(self.ptr, C.byref(c), C.byref(r)))
        return c.value, r.value

    def batch_isects(self) -> int:
        """Tile intersections of the last batched frame on this context."""
        m, v = C.c_uint64(), C.c_uint64()
        check(lib().msplat_hip_batch_frame_info(self.ptr, C.byref(m), C.byref(v)))
        return m.value

    def download(self, ptr, shape, dtype) -> np.ndarray:
        out = np.empty(shape, dtype)
        if out.nbytes:
            check(lib().msplat_hip_memcpy(self.ptr, out.ctypes.data, C.c_void_p(ptr), out.nbytes, 1))
        return out

    def upload(self, ptr, arr: np.ndarray):
        arr = np.ascontiguousarray(arr)
        if arr.nbytes:
            check(lib().msplat_hip_memcpy(self.ptr, C.c_void_p(ptr), arr.ctypes.data, arr.nbytes, 0))


class DeviceArray:
    """One device allocation holding a numpy-shaped array."""

    def __init__(self, ctx: Context, shape, dtype=np.float32, zero=True):
        self.ctx, self.shape, self.dtype = ctx, tuple(shape), np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        check(lib().msplat_hip_malloc(max(self.nbytes, 4), 1 if zero else 0, C.byref(p)))
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, ctx, arr, dtype=None):
        arr = np.ascontiguousarray(arr, dtype=dtype or np.asarray(arr).dtype)
        d = cls(ctx, arr.shape, arr.dtype, zero=False)
        d.upload(arr)
        return d

    def upload(self, arr):
        arr = np.ascontiguousarray(arr, dtype=self.dtype).reshape(self.shape)
        self.ctx.upload(self.ptr, arr)

    def numpy(self) -> np.ndarray:
        return self.ctx.download(self.ptr, self.shape, self.dtype)

    def free(self):
        if self.ptr:
            lib().msplat_hip_free(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_memory(device: int = 0) -> tuple:
    """(free, total) bytes of device memory."""
    free, total = C.c_size_t(), C.c_size_t()
    check(lib().msplat_hip_device_memory(device, C.byref(free), C.byref(total)))
    return free.value, total.value


def make_camera_struct(cam: dict, degree: int, glob_scale: float = 1.0, clip: float = 0.01) -> Camera:
    c = Camera()
    c.viewmat[:] = [float(v) for v in np.asarray(cam["viewmat"], np.float32).ravel()]
    c.projmat[:] = [float(v) for v in np.asarray(cam["projmat"], np.float32).ravel()]
    c.cam_pos[:] = [float(v) for v in np.asarray(cam["cam_pos"], np.float32).ravel()]
    c.fx, c.fy, c.cx, c.cy = (float(cam[k]) for k in ("fx", "fy", "cx", "cy"))
    c.width, c.height = int(cam["W"]), int(cam["H"])
    c.tiles_x, c.tiles_y = int(cam["tiles_x"]), int(cam["tiles_y"])
    c.clip_thresh, c.glob_scale = clip, glob_scale
    c.degree, c.degrees_to_use = int(degree), int(cam["dtu"])
    return c


def floats(v) -> "C.Array":
    a = np.ascontiguousarray(v, np.float32)
    return (C.c_float * a.size)(*a.tolist())
