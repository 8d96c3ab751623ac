"""ctypes wrappers over the oracle (liboracle.so) and the compiled reference
(_ref/libmsplat_ref*.so). TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check
and the CPU-baseline parts of bench.py (skipped by --no-cpu-baseline), never by the
product.

The oracle restates core/cpu/msplat_cpu.cpp (see msplat_oracle.cpp for the per-function
citations); `Reference` calls the reference's own CPU backend through ref_driver.cpp.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
vp = C.c_void_p


def _load(name):
    path = os.path.join(_HERE, name)
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is not built (make -C oracle)")
    return C.CDLL(path)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def _ptrs(arrays):
    return (C.c_void_p * len(arrays))(*[a.ctypes.data for a in arrays])


def sh_bases(degree: int) -> int:
    return {0: 1, 1: 4, 2: 9, 3: 16}.get(degree, 25)


class PriorStruct(C.Structure):
    """PriorStep (core/metal/bindings.h:43-77) as plain pointers: the oracle's Prior and
    ref_driver.cpp's RefPrior."""
    _fields_ = [("aux", C.c_int32), ("sky", C.c_int32), ("sky_tex", vp), ("sky_grad", vp), ("sky_m1", vp),
                ("sky_m2", vp), ("sky_w", C.c_int32), ("sky_h", C.c_int32), ("sky_frame", C.c_float * 9),
                ("sky_step", C.c_float), ("sky_bc2", C.c_float), ("prior_depth", vp), ("prior_aux", vp),
                ("prior_w", C.c_int32), ("prior_h", C.c_int32), ("has_depth", C.c_int32), ("has_sky_mask", C.c_int32),
                ("depth_weight", C.c_float), ("sky_weight", C.c_float), ("fill_weight", C.c_float),
                ("huber_delta", C.c_float), ("min_alpha", C.c_float), ("mask_photometric", C.c_int32),
                ("loss_terms", vp), ("exposure", C.c_int32), ("cam_index", C.c_int32), ("num_cams", C.c_int32),
                ("expo_params", vp), ("expo_m1", vp), ("expo_m2", vp), ("expo_grad", vp), ("expo_step", C.c_float),
                ("expo_bc2", C.c_float), ("expo_reg", C.c_float), ("log_max_scale_ratio", C.c_float)]


PRIOR_ARRAYS = {"sky_tex": np.float32, "sky_grad": np.float32, "sky_m1": np.float32, "sky_m2": np.float32,
                "prior_depth": np.float32, "prior_aux": np.uint8, "loss_terms": np.float32, "expo_params": np.float32,
                "expo_m1": np.float32, "expo_m2": np.float32, "expo_grad": np.float32}
PRIOR_DEFAULTS = dict(aux=False, sky=False, sky_frame=(0.0,) * 9, sky_step=0.0, sky_bc2=1.0, has_depth=False,
                      has_sky_mask=False, depth_weight=0.0, sky_weight=0.0, fill_weight=0.0, huber_delta=0.05,
                      min_alpha=0.25, mask_photometric=True, exposure=False, cam_index=0, expo_step=0.0, expo_bc2=1.0,
                      expo_reg=0.0, log_max_scale_ratio=0.0)


PRIOR_INTS = ("aux", "sky", "has_depth", "has_sky_mask", "mask_photometric", "exposure", "cam_index")


def prior_struct(prior: dict) -> PriorStruct:
    """The C record of a prior given as a dict of PriorStep's fields (numpy arrays for the
    tensors, which must be C-contiguous arrays of the listed dtypes: the C side updates
    them in place)."""
    st = PriorStruct()
    for k, v in PRIOR_DEFAULTS.items():
        v = prior.get(k, v)
        if k == "sky_frame":
            st.sky_frame[:] = [float(x) for x in v]
        else:
            setattr(st, k, int(v) if k in PRIOR_INTS else float(v))
    for k, dt in PRIOR_ARRAYS.items():
        a = prior.get(k)
        if a is None:
            continue
        if a.dtype != dt or not a.flags["C_CONTIGUOUS"]:
            raise ValueError(f"prior[{k!r}] must be a C-contiguous {np.dtype(dt).name} array")
        setattr(st, k, a.ctypes.data)
    if prior.get("sky_tex") is not None:
        st.sky_h, st.sky_w = prior["sky_tex"].shape[:2]
    if prior.get("prior_depth") is not None:
        st.prior_h, st.prior_w = prior["prior_depth"].shape[:2]
    if prior.get("expo_params") is not None:
        st.num_cams = prior["expo_params"].shape[0]
    return st


class Oracle:
    """The CPU restatement. Methods take/return numpy arrays in the reference layouts."""

    def __init__(self):
        L = self.lib = _load("liboracle.so")
        for s, fp in (("f32", f32p), ("f64", f64p)):
            T = C.c_float if s == "f32" else C.c_double
            f = getattr(L, f"orc_{s}_project")
            f.argtypes = [C.c_int, fp, fp, T, fp, fp, fp, T, T, T, T, C.c_uint32, C.c_uint32, C.c_int, C.c_int, T,
                          C.c_uint32, C.c_uint32, fp, fp, fp, fp, fp, fp, fp, fp, i32p]
            f = getattr(L, f"orc_{s}_count_isects")
            f.argtypes = [C.c_int, fp, fp, i32p, C.c_int, C.c_int, vp]
            f.restype = C.c_int64
            f = getattr(L, f"orc_{s}_bin_sort")
            f.argtypes = [C.c_int, fp, fp, fp, i32p, fp, fp, fp, C.c_int, C.c_int, i32p, i32p, i32p, fp, fp, fp, vp]
            f = getattr(L, f"orc_{s}_raster_forward")
            f.argtypes = [C.c_uint32, C.c_uint32, C.c_int, C.c_int, i32p, i32p, fp, fp, fp, fp, fp, fp, i32p]
            f = getattr(L, f"orc_{s}_raster_backward")
            f.argtypes = [C.c_uint32, C.c_uint32, C.c_int, C.c_int, i32p, i32p, fp, fp, fp, fp, fp, i32p, fp, fp]
            f = getattr(L, f"orc_{s}_reduce_slots")
            f.argtypes = [C.c_int, C.c_int64, i32p, fp, fp, fp, fp, fp, fp]
        L.orc_f32_loss.argtypes = [C.c_uint32, C.c_uint32, f32p, f32p, C.c_float, C.c_float, f32p]
        L.orc_f32_loss.restype = C.c_double
        L.orc_f32_project_backward_adam.argtypes = (
            [C.c_int, f32p, f32p, C.c_float, f32p, f32p, f32p, C.c_float, C.c_float, C.c_uint32, C.c_uint32,
             C.c_uint32, C.c_uint32, f32p, i32p, f32p, f32p, f32p, f32p, f32p, f32p, f32p, f32p, f32p, f32p, f32p,
             f32p, f32p, f32p] + [C.c_float] * 7)
        L.orc_f64_project_backward_adam.argtypes = (
            [C.c_int, f64p, f64p, C.c_double, f64p, f64p, f64p, C.c_double, C.c_double, C.c_uint32, C.c_uint32,
             C.c_uint32, C.c_uint32, f64p, i32p, f64p, f64p, f64p, f64p, f64p, f64p, f64p, f64p, f64p, f64p, f64p,
             f64p, f64p, f64p] + [C.c_double] * 7)
        L.orc_f32_adam.argtypes = [f32p, f32p, f32p, f32p, C.c_size_t] + [C.c_float] * 5
        L.orc_f32_grad_stats.argtypes = [C.c_int, i32p, f32p, C.c_float, f32p, f32p, f32p]
        L.orc_f32_train_step.argtypes = (
            [C.c_int, C.c_float, f32p, f32p, C.c_float, C.c_float, C.c_float, C.c_float, C.c_uint32, C.c_uint32,
             C.c_int, C.c_int, C.c_float, C.c_uint32, C.c_uint32, f32p, f32p, f32p, C.c_float, C.c_float, vp, vp, vp,
             f32p, f32p, C.c_float, C.c_float, C.c_float, f32p, f32p, f32p, C.c_float, vp, vp])
        L.orc_f32_train_step.restype = C.c_float
        L.orc_f32_train_step_prior.argtypes = L.orc_f32_train_step.argtypes + [C.POINTER(PriorStruct), vp, vp]
        L.orc_f32_train_step_prior.restype = C.c_float
        L.orc_f32_render_prior.argtypes = (
            [C.c_int, C.c_float, f32p, f32p, C.c_float, C.c_float, C.c_float, C.c_float, C.c_uint32, C.c_uint32,
             C.c_int, C.c_int, C.c_float, C.c_uint32, C.c_uint32, f32p, vp, f32p, C.POINTER(PriorStruct), f32p, vp, vp,
             vp])
        L.orc_prepare_cam.argtypes = [f32p, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int,
                                      f32p, f32p, f32p, f32p, i32p]
        L.orc_means_lr.argtypes = [C.c_int, C.c_int, C.c_float, C.c_float]
        L.orc_means_lr.restype = C.c_float
        L.orc_adam_scalars.argtypes = [C.c_float, C.c_float, C.c_float, C.c_int, C.POINTER(C.c_float),
                                       C.POINTER(C.c_float)]
        L.orc_refine_flags.argtypes = [C.c_int] * 7
        L.orc_refine_flags.restype = C.c_int
        L.orc_densify.argtypes = ([C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_int, C.c_float, C.c_float,
                                   C.c_float, C.c_int, f32p, f32p, f32p, C.c_float, vp, C.c_int, C.c_uint32, vp, vp,
                                   vp, vp])
        L.orc_opacity_reset.argtypes = [f32p, f32p, f32p, C.c_int, C.c_float]
        L.orc_radix_sort.argtypes = [u64p, i32p, C.c_uint32, C.c_int]
        L.orc_knn3_mean_dist.argtypes = [f32p, C.c_uint32, f32p]
        L.orc_init_gaussians.argtypes = [f32p, u8p, C.c_uint32, C.c_uint32, C.c_float, f32p, f32p, f32p, f32p]
        L.orc_image_metrics.argtypes = [f32p, f32p, C.c_int, C.c_int, f64p]
        L.orc_pack_rgba8.argtypes = [f32p, C.c_uint32, u8p]
        L.orc_resize_area.argtypes = [vp, C.c_int, C.c_int, C.c_int, f32p, C.c_int, C.c_int]
        L.orc_undistort.argtypes = [f32p, C.c_int, C.c_int, f32p, f32p, C.c_int, C.c_int, f32p, C.c_int, C.c_int]
        L.orc_finalize_depth.argtypes = [f32p, f32p, C.c_uint32, C.c_float, f32p, f32p]
        L.orc_pcg_hash.argtypes = [C.c_uint32]
        L.orc_pcg_hash.restype = C.c_uint32
        L.orc_rand_uniform.argtypes = [C.c_uint32, C.c_uint32]
        L.orc_rand_uniform.restype = C.c_float
        L.orc_rand_normal3.argtypes = [C.c_uint32, C.c_uint32, f32p]

    # ── render path (float32 or float64) ────────────────────────────────────
    def project(self, scene, cam, dtype=np.float32, glob_scale=1.0, clip=0.01):
        s = "f32" if dtype == np.float32 else "f64"
        N = scene["means"].shape[0]
        out = dict(xys=np.zeros((N, 2), dtype), depths=np.zeros(N, dtype), conics=np.zeros((N, 3), dtype),
                   colors=np.zeros((N, 3), dtype), aabb=np.zeros((N, 2), dtype), radii=np.zeros(N, np.int32))
        getattr(self.lib, f"orc_{s}_project")(
            N, _c(scene["means"], dtype), _c(scene["scales"], dtype), glob_scale, _c(scene["quats"], dtype),
            _c(cam["viewmat"], dtype), _c(cam["projmat"], dtype), cam["fx"], cam["fy"], cam["cx"], cam["cy"],
            cam["W"], cam["H"], cam["tiles_x"], cam["tiles_y"], clip, scene["degree"], cam["dtu"],
            _c(cam["cam_pos"], dtype), _c(scene["dc"], dtype), _c(scene["rest"], dtype), out["xys"], out["depths"],
            out["conics"], out["colors"], out["aabb"], out["radii"])
        return out

    def bin_sort(self, proj, opac, cam, dtype=np.float32):
        s = "f32" if dtype == np.float32 else "f64"
        N = proj["radii"].shape[0]
        tx, ty = cam["tiles_x"], cam["tiles_y"]
        args = (_c(proj["xys"], dtype), _c(proj["depths"], dtype), _c(proj["aabb"], dtype), _c(proj["radii"], np.int32))
        M = getattr(self.lib, f"orc_{s}_count_isects")(N, args[0], args[2], args[3], tx, ty, None)
        Mc = max(M, 1)
        out = dict(tile_start=np.zeros(tx * ty, np.int32), tile_end=np.zeros(tx * ty, np.int32),
                   isect_gid=np.zeros(Mc, np.int32), pk_xyo=np.zeros((Mc, 3), dtype), pk_conic=np.zeros((Mc, 3), dtype),
                   pk_rgb=np.zeros((Mc, 3), dtype), keys=np.zeros(Mc, np.uint64), M=M)
        getattr(self.lib, f"orc_{s}_bin_sort")(
            N, *args[:2], args[2], args[3], _c(np.ravel(opac), dtype), _c(proj["conics"], dtype),
            _c(proj["colors"], dtype), tx, ty, out["tile_start"], out["tile_end"], out["isect_gid"], out["pk_xyo"],
            out["pk_conic"], out["pk_rgb"], out["keys"].ctypes.data)
        for k in ("isect_gid", "pk_xyo", "pk_conic", "pk_rgb", "keys"):
            out[k] = out[k][:M]
        return out

    def raster_forward(self, bins, cam, bg, dtype=np.float32):
        s = "f32" if dtype == np.float32 else "f64"
        W, H = cam["W"], cam["H"]
        img = np.zeros((H, W, 3), dtype)
        fT = np.zeros((H, W), dtype)
        fidx = np.zeros((H, W), np.int32)
        Mc = max(bins["M"], 1)
        getattr(self.lib, f"orc_{s}_raster_forward")(
            W, H, cam["tiles_x"], cam["tiles_y"], bins["tile_start"], bins["tile_end"],
            _pad(bins["pk_xyo"], Mc, dtype), _pad(bins["pk_conic"], Mc, dtype), _pad(bins["pk_rgb"], Mc, dtype),
            _c(bg, dtype), img, fT, fidx)
        return dict(img=img, final_T=fT, final_idx=fidx)

    def render(self, scene, cam, bg, dtype=np.float32):
        p = self.project(scene, cam, dtype)
        b = self.bin_sort(p, scene["opac"], cam, dtype)
        r = self.raster_forward(b, cam, bg, dtype)
        return p, b, r

    def loss(self, img, gt, ssim_weight=0.2):
        H, W = img.shape[:2]
        v = np.zeros((H, W, 3), np.float32)
        tot = self.lib.orc_f32_loss(W, H, _c(img, np.float32), _c(gt, np.float32), ssim_weight,
                                    np.float32(1.0) / np.float32(H * W * 3), v)
        return tot, v

    def raster_backward(self, bins, fwd, cam, bg, v_out, dtype=np.float32):
        s = "f32" if dtype == np.float32 else "f64"
        Mc = max(bins["M"], 1)
        slots = np.zeros((Mc, 10), dtype)
        getattr(self.lib, f"orc_{s}_raster_backward")(
            cam["W"], cam["H"], cam["tiles_x"], cam["tiles_y"], bins["tile_start"], bins["tile_end"],
            _pad(bins["pk_xyo"], Mc, dtype), _pad(bins["pk_conic"], Mc, dtype), _pad(bins["pk_rgb"], Mc, dtype),
            _c(bg, dtype), _c(fwd["final_T"], dtype), _c(fwd["final_idx"], np.int32), _c(v_out, dtype), slots)
        return slots[:bins["M"]]

    def reduce_slots(self, N, bins, slots, dtype=np.float32):
        s = "f32" if dtype == np.float32 else "f64"
        Mc = max(bins["M"], 1)
        out = dict(v_xy=np.zeros((N, 2), dtype), v_conic=np.zeros((N, 3), dtype), v_rgb=np.zeros((N, 3), dtype),
                   v_opac=np.zeros(N, dtype), v_depth=np.zeros(N, dtype))
        getattr(self.lib, f"orc_{s}_reduce_slots")(
            N, bins["M"], _pad(bins["isect_gid"], Mc, np.int32), _pad(slots, Mc, dtype), out["v_xy"], out["v_conic"],
            out["v_rgb"], out["v_opac"], out["v_depth"])
        return out

    def gradients(self, scene, cam, bg, v_img, dtype=np.float32):
        """Image, final transmittance and the parameter gradients of <image, v_img> by the
        analytic backward (raster backward, per-Gaussian reduction, projection VJPs) in
        float32 or float64: {img, final_T, means, scales, quats, dc, rest, opac}."""
        s = "f32" if dtype == np.float32 else "f64"
        p, b, r = self.render(scene, cam, bg, dtype)
        N = scene["means"].shape[0]
        slots = self.raster_backward(b, r, cam, bg, v_img, dtype)
        red = self.reduce_slots(N, b, slots, dtype)
        nb = sh_bases(scene["degree"])
        out = dict(img=r["img"], final_T=r["final_T"], means=np.zeros((N, 3), dtype), scales=np.zeros((N, 3), dtype),
                   quats=np.zeros((N, 4), dtype), opac=red["v_opac"].reshape(N, 1).copy())
        dc, rest = _c(scene["dc"], dtype).copy(), _c(scene["rest"], dtype).copy()
        g_dc, g_rest = np.zeros_like(dc), np.zeros_like(rest)
        m2_dc, m2_rest = np.zeros_like(dc), np.zeros_like(rest)
        getattr(self.lib, f"orc_{s}_project_backward_adam")(
            N, _c(scene["means"], dtype), _c(scene["scales"], dtype), 1.0, _c(scene["quats"], dtype),
            _c(cam["viewmat"], dtype), _c(cam["projmat"], dtype), cam["fx"], cam["fy"], cam["W"], cam["H"],
            scene["degree"], cam["dtu"], _c(cam["cam_pos"], dtype), p["radii"], p["conics"], red["v_xy"],
            red["v_conic"], red["v_rgb"], red["v_depth"], out["means"], out["scales"], out["quats"], dc, rest, g_dc,
            m2_dc, g_rest, m2_rest, 0.0, 1.0, 0.0, 1.0, 0.0, 0.0, 1e-8)
        out["dc"], out["rest"] = g_dc, g_rest.reshape(N, max(nb - 1, 1), 3)
        return out

    def train_step(self, state, cam, bg, gt, steps, bc2s, ssim_weight=0.2, b1=0.9, b2=0.999, eps=1e-8):
        """One training step in place on `state` (params p0..p5, m1_*, m2_*, vis/gn/ms).
        Returns (loss, radii, image)."""
        N = state["means"].shape[0]
        names = ("means", "scales", "quats", "dc", "rest", "opac")
        for n in names:
            for pre in ("", "m1_", "m2_"):
                state[pre + n] = _c(state[pre + n], np.float32)
        radii = np.zeros(N, np.int32)
        img = np.zeros((cam["H"], cam["W"], 3), np.float32)
        H, W = cam["H"], cam["W"]
        loss = self.lib.orc_f32_train_step(
            N, 1.0, _c(cam["viewmat"], np.float32), _c(cam["projmat"], np.float32), cam["fx"], cam["fy"], cam["cx"],
            cam["cy"], H, W, cam["tiles_x"], cam["tiles_y"], 0.01, state["degree"], cam["dtu"],
            _c(cam["cam_pos"], np.float32), _c(bg, np.float32), _c(gt, np.float32), ssim_weight,
            np.float32(1.0) / np.float32(H * W * 3), _ptrs([state[n] for n in names]),
            _ptrs([state["m1_" + n] for n in names]), _ptrs([state["m2_" + n] for n in names]),
            _c(steps, np.float32), _c(bc2s, np.float32), b1, b2, eps, state["vis"], state["gn"], state["ms"],
            np.float32(1.0) / np.float32(max(H, W)), radii.ctypes.data, img.ctypes.data)
        return loss, radii, img

    def train_step_prior(self, state, cam, bg, gt, steps, bc2s, prior: dict, ssim_weight=0.2, b1=0.9, b2=0.999,
                         eps=1e-8):
        """msplat_train_step with a PriorStep (given as a dict, see prior_struct), in place
        on `state` and on the prior's arrays. Returns (loss, radii, image, depth, final_T)."""
        N = state["means"].shape[0]
        names = ("means", "scales", "quats", "dc", "rest", "opac")
        for n in names:
            for pre in ("", "m1_", "m2_"):
                state[pre + n] = _c(state[pre + n], np.float32)
        H, W = cam["H"], cam["W"]
        radii = np.zeros(N, np.int32)
        img = np.zeros((H, W, 3), np.float32)
        depth = np.zeros((H, W), np.float32)
        fT = np.zeros((H, W), np.float32)
        st = prior_struct(prior)
        loss = self.lib.orc_f32_train_step_prior(
            N, 1.0, _c(cam["viewmat"], np.float32), _c(cam["projmat"], np.float32), cam["fx"], cam["fy"], cam["cx"],
            cam["cy"], H, W, cam["tiles_x"], cam["tiles_y"], 0.01, state["degree"], cam["dtu"],
            _c(cam["cam_pos"], np.float32), _c(bg, np.float32), _c(gt, np.float32), ssim_weight,
            np.float32(1.0) / np.float32(H * W * 3), _ptrs([state[n] for n in names]),
            _ptrs([state["m1_" + n] for n in names]), _ptrs([state["m2_" + n] for n in names]),
            _c(steps, np.float32), _c(bc2s, np.float32), b1, b2, eps, state["vis"], state["gn"], state["ms"],
            np.float32(1.0) / np.float32(max(H, W)), radii.ctypes.data, img.ctypes.data, C.byref(st),
            depth.ctypes.data, fT.ctypes.data)
        return loss, radii, img, depth, fT

    def render_prior(self, scene, cam, bg, prior: dict):
        """msplat_render with a PriorStep: (image, expected-depth sum, final transmittance, radii)."""
        N = scene["means"].shape[0]
        H, W = cam["H"], cam["W"]
        img = np.zeros((H, W, 3), np.float32)
        depth = np.zeros((H, W), np.float32)
        fT = np.zeros((H, W), np.float32)
        radii = np.zeros(N, np.int32)
        params = [_c(scene[k], np.float32) for k in ("means", "scales", "quats", "dc", "rest", "opac")]
        st = prior_struct(prior)
        self.lib.orc_f32_render_prior(
            N, 1.0, _c(cam["viewmat"], np.float32), _c(cam["projmat"], np.float32), cam["fx"], cam["fy"], cam["cx"],
            cam["cy"], H, W, cam["tiles_x"], cam["tiles_y"], 0.01, scene["degree"], cam["dtu"],
            _c(cam["cam_pos"], np.float32), _ptrs(params), _c(bg, np.float32), C.byref(st), img, depth.ctypes.data,
            fT.ctypes.data, radii.ctypes.data)
        return img, depth, fT, radii

    def densify(self, bufs, N, cap, gn, vc, ms, half_max_dim, fr_stride, seed, grad_thr=2e-4, size_thr=0.01,
                screen_thr=0.05, check_screen=1, cull_alpha=0.1, cull_scale=0.5, cull_screen=0.15, check_huge=0):
        """bufs: list of 18 float32 arrays with `cap` rows, modified in place. Returns
        (new_count, split, dup, keep)."""
        split = np.zeros(max(N, 1), np.uint8)
        dup = np.zeros(max(N, 1), np.uint8)
        keep = np.zeros(max(3 * N, 1), np.uint8)
        nn = C.c_int(0)
        out = self.lib.orc_densify(N, cap, grad_thr, size_thr, screen_thr, check_screen, cull_alpha, cull_scale,
                                   cull_screen, check_huge, _c(gn, np.float32), _c(vc, np.float32),
                                   _c(ms, np.float32), half_max_dim, _ptrs(bufs), fr_stride, seed,
                                   split.ctypes.data, dup.ctypes.data, keep.ctypes.data, C.byref(nn))
        return out, split[:N], dup[:N], keep[:nn.value]

    def prepare_cam(self, cam_to_world, W, H, fx, fy, cx, cy, sf=1):
        """Model::prepareCam's matrices, intrinsics and size at downscale factor sf."""
        vm, pm, pos, intr = (np.zeros(k, np.float32) for k in (16, 16, 3, 4))
        wh = np.zeros(2, np.int32)
        self.lib.orc_prepare_cam(_c(cam_to_world, np.float32).reshape(16), int(W), int(H), fx, fy, cx, cy, int(sf),
                                 vm, pm, pos, intr, wh)
        return dict(viewmat=vm, projmat=pm, cam_pos=pos, fx=intr[0], fy=intr[1], cx=intr[2], cy=intr[3],
                    W=int(wh[0]), H=int(wh[1]))

    def means_lr(self, step, max_steps, lr_init=1.6e-4, lr_final=1.6e-6):
        return np.float32(self.lib.orc_means_lr(int(step), int(max_steps), lr_init, lr_final))

    def adam_scalars(self, lr, t, beta1=0.9, beta2=0.999):
        a, b = C.c_float(), C.c_float()
        self.lib.orc_adam_scalars(lr, beta1, beta2, int(t), C.byref(a), C.byref(b))
        return np.float32(a.value), np.float32(b.value)

    def refine_flags(self, step, refine_every, warmup, reset_alpha_every, stop_split_at, num_cameras,
                     stop_screen_size_at):
        return int(self.lib.orc_refine_flags(step, refine_every, warmup, reset_alpha_every, stop_split_at,
                                             num_cameras, stop_screen_size_at))

    def opacity_reset(self, o, m1, m2, reset_logit):
        self.lib.orc_opacity_reset(o, m1, m2, o.size, reset_logit)

    def radix_sort(self, keys, vals, key_bits):
        k = _c(keys, np.uint64).copy()
        v = _c(vals, np.int32).copy()
        self.lib.orc_radix_sort(k, v, k.size, key_bits)
        return k, v

    def knn3(self, pts):
        pts = _c(pts, np.float32)
        out = np.zeros(pts.shape[0], np.float32)
        self.lib.orc_knn3_mean_dist(pts, pts.shape[0], out)
        return out

    def init_gaussians(self, means, rgb, seed=42, logit=None):
        n = means.shape[0]
        logit = np.float32(np.log(np.float32(0.1) / np.float32(0.9))) if logit is None else logit
        s, q = np.zeros((n, 3), np.float32), np.zeros((n, 4), np.float32)
        d, o = np.zeros((n, 3), np.float32), np.zeros((n, 1), np.float32)
        self.lib.orc_init_gaussians(_c(means, np.float32), _c(rgb, np.uint8), n, seed, logit, s, q, d, o)
        return s, q, d, o

    def image_metrics(self, a, b):
        out = np.zeros(3, np.float64)
        self.lib.orc_image_metrics(_c(a, np.float32), _c(b, np.float32), a.shape[0], a.shape[1], out)
        return out

    def pack_rgba8(self, img):
        n = img.shape[0] * img.shape[1]
        out = np.zeros((img.shape[0], img.shape[1], 4), np.uint8)
        self.lib.orc_pack_rgba8(_c(img, np.float32), n, out)
        return out

    def resize_area(self, src, dw, dh):
        """src: (sh, sw, 4) uint8 RGBA or (sh, sw, 3) float32 -> (dh, dw, 3) float32."""
        src = np.ascontiguousarray(src)
        out = np.zeros((dh, dw, 3), np.float32)
        self.lib.orc_resize_area(src.ctypes.data, int(src.dtype == np.uint8), src.shape[1], src.shape[0], out, dw, dh)
        return out

    def undistort(self, src, intr, dist, roi_x, roi_y, dw, dh):
        out = np.zeros((dh, dw, 3), np.float32)
        self.lib.orc_undistort(_c(src, np.float32), src.shape[1], src.shape[0], _c(intr, np.float32),
                               _c(dist, np.float32), roi_x, roi_y, out, dw, dh)
        return out

    def finalize_depth(self, depth_num, final_T, inv_scale=1.0):
        d, t = _c(depth_num, np.float32).ravel(), _c(final_T, np.float32).ravel()
        od, oa = np.zeros_like(d), np.zeros_like(d)
        self.lib.orc_finalize_depth(d, t, d.size, inv_scale, od, oa)
        return od.reshape(np.shape(depth_num)), oa.reshape(np.shape(depth_num))

    def pcg_hash(self, v):
        return self.lib.orc_pcg_hash(v)

    def rand_uniform(self, seed, ctr):
        return self.lib.orc_rand_uniform(seed, ctr)

    def rand_normal3(self, seed, stream):
        out = np.zeros(3, np.float32)
        self.lib.orc_rand_normal3(seed, stream, out)
        return out


def _pad(a, n, dtype):
    a = np.ascontiguousarray(a, dtype=dtype)
    if a.shape[0] >= n:
        return a
    pad = np.zeros((n - a.shape[0],) + a.shape[1:], dtype)
    return np.concatenate([a, pad])


class Reference:
    """The reference's compiled CPU backend (core/cpu/msplat_cpu.cpp), built from the
    mounted reference tree into oracle/_ref by oracle/Makefile."""

    def __init__(self, perf: bool = False, isa: str = "v3"):
        name = "libmsplat_ref.so" if not perf else ("libmsplat_ref_perf_v4.so" if isa == "v4" else
                                                     "libmsplat_ref_perf.so")
        L = self.lib = _load(os.path.join("_ref", name))
        L.ref_render.argtypes = ([C.c_int, f32p, f32p, f32p, f32p, f32p, C.c_int, f32p, f32p, f32p, C.c_float,
                                  C.c_float, C.c_float, C.c_float, C.c_int, C.c_int, C.c_float, C.c_int, C.c_int,
                                  f32p, f32p, f32p, vp])
        L.ref_train_step.argtypes = ([C.c_int, vp, C.c_int, vp, vp, f32p, f32p, C.c_float, C.c_float, C.c_float,
                                      C.c_float, C.c_int, C.c_int, C.c_float, C.c_int, C.c_int, f32p, f32p, f32p,
                                      C.c_float, f32p, f32p, C.c_float, C.c_float, C.c_float, f32p, f32p, f32p, i32p,
                                      vp])
        L.ref_densify.argtypes = ([C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_int, C.c_float, C.c_float,
                                   C.c_float, C.c_int, f32p, f32p, f32p, C.c_float, vp, C.c_int, C.c_uint32])
        L.ref_train_step_prior.argtypes = L.ref_train_step.argtypes + [C.POINTER(PriorStruct), vp, vp]
        L.ref_finalize_depth.argtypes = [f32p, f32p, C.c_uint32, C.c_float, f32p, f32p]
        L.ref_render_prior.argtypes = ([C.c_int, f32p, f32p, f32p, f32p, f32p, C.c_int, f32p, f32p, f32p, C.c_float,
                                        C.c_float, C.c_float, C.c_float, C.c_int, C.c_int, C.c_float, C.c_int,
                                        C.c_int, f32p, f32p, C.POINTER(PriorStruct), f32p, vp, vp])
        L.ref_opacity_reset.argtypes = [f32p, f32p, f32p, C.c_int, C.c_float]
        L.ref_radix_sort.argtypes = [u64p, i32p, C.c_uint32, C.c_int]
        L.ref_knn3.argtypes = [f32p, C.c_uint32, f32p]
        L.ref_init_gaussians.argtypes = [f32p, u8p, C.c_uint32, C.c_uint32, C.c_float, f32p, f32p, f32p, f32p]
        L.ref_image_metrics.argtypes = [f32p, f32p, C.c_int, C.c_int, f64p]
        L.ref_pack_rgba8.argtypes = [f32p, C.c_uint32, u8p]
        L.ref_resize_area.argtypes = [vp, C.c_int, C.c_int, C.c_int, f32p, C.c_int, C.c_int]
        L.ref_undistort.argtypes = [f32p, C.c_int, C.c_int, f32p, f32p, C.c_int, C.c_int, f32p, C.c_int, C.c_int]

    def render(self, scene, cam, bg):
        N = scene["means"].shape[0]
        img = np.zeros((cam["H"], cam["W"], 3), np.float32)
        fT = np.zeros((cam["H"], cam["W"]), np.float32)
        rb = sh_bases(scene["degree"]) - 1
        rc = self.lib.ref_render(
            N, _c(scene["means"], np.float32), _c(scene["scales"], np.float32), _c(scene["quats"], np.float32),
            _c(scene["dc"], np.float32), _c(scene["rest"], np.float32), rb, _c(scene["opac"], np.float32),
            _c(cam["viewmat"], np.float32), _c(cam["projmat"], np.float32), cam["fx"], cam["fy"], cam["cx"],
            cam["cy"], cam["H"], cam["W"], 0.01, scene["degree"], cam["dtu"], _c(cam["cam_pos"], np.float32),
            _c(bg, np.float32), img, fT.ctypes.data)
        if rc:
            raise RuntimeError("reference render failed")
        return img, fT

    def train_step(self, state, cam, bg, gt, steps, bc2s, ssim_weight=0.2, b1=0.9, b2=0.999, eps=1e-8):
        N = state["means"].shape[0]
        names = ("means", "scales", "quats", "dc", "rest", "opac")
        for n in names:
            for pre in ("", "m1_", "m2_"):
                state[pre + n] = _c(state[pre + n], np.float32)
        radii = np.zeros(N, np.int32)
        loss = C.c_float(0)
        rc = self.lib.ref_train_step(
            N, _ptrs([state[n] for n in names]), sh_bases(state["degree"]) - 1,
            _ptrs([state["m1_" + n] for n in names]), _ptrs([state["m2_" + n] for n in names]),
            _c(cam["viewmat"], np.float32), _c(cam["projmat"], np.float32), cam["fx"], cam["fy"], cam["cx"],
            cam["cy"], cam["H"], cam["W"], 0.01, state["degree"], cam["dtu"], _c(cam["cam_pos"], np.float32),
            _c(bg, np.float32), _c(gt, np.float32), ssim_weight, _c(steps, np.float32), _c(bc2s, np.float32), b1, b2,
            eps, state["vis"], state["gn"], state["ms"], radii, C.byref(loss))
        if rc:
            raise RuntimeError("reference train_step failed")
        return loss.value, radii

    def train_step_prior(self, state, cam, bg, gt, steps, bc2s, prior: dict, ssim_weight=0.2, b1=0.9, b2=0.999,
                         eps=1e-8):
        N = state["means"].shape[0]
        names = ("means", "scales", "quats", "dc", "rest", "opac")
        for n in names:
            for pre in ("", "m1_", "m2_"):
                state[pre + n] = _c(state[pre + n], np.float32)
        H, W = cam["H"], cam["W"]
        radii = np.zeros(N, np.int32)
        depth = np.zeros((H, W), np.float32)
        fT = np.zeros((H, W), np.float32)
        loss = C.c_float(0)
        st = prior_struct(prior)
        rc = self.lib.ref_train_step_prior(
            N, _ptrs([state[n] for n in names]), sh_bases(state["degree"]) - 1,
            _ptrs([state["m1_" + n] for n in names]), _ptrs([state["m2_" + n] for n in names]),
            _c(cam["viewmat"], np.float32), _c(cam["projmat"], np.float32), cam["fx"], cam["fy"], cam["cx"],
            cam["cy"], H, W, 0.01, state["degree"], cam["dtu"], _c(cam["cam_pos"], np.float32),
            _c(bg, np.float32), _c(gt, np.float32), ssim_weight, _c(steps, np.float32), _c(bc2s, np.float32), b1, b2,
            eps, state["vis"], state["gn"], state["ms"], radii, C.byref(loss), C.byref(st), depth.ctypes.data,
            fT.ctypes.data)
        if rc:
            raise RuntimeError("reference train_step with a prior failed")
        return loss.value, radii, depth, fT

    def finalize_depth(self, depth_num, final_T, inv_scale=1.0):
        d, t = _c(depth_num, np.float32).ravel(), _c(final_T, np.float32).ravel()
        od, oa = np.zeros_like(d), np.zeros_like(d)
        if self.lib.ref_finalize_depth(d, t, d.size, inv_scale, od, oa):
            raise RuntimeError("reference finalize_depth failed")
        return od.reshape(np.shape(depth_num)), oa.reshape(np.shape(depth_num))

    def render_prior(self, scene, cam, bg, prior: dict):
        N = scene["means"].shape[0]
        H, W = cam["H"], cam["W"]
        img = np.zeros((H, W, 3), np.float32)
        depth = np.zeros((H, W), np.float32)
        fT = np.zeros((H, W), np.float32)
        rb = sh_bases(scene["degree"]) - 1
        st = prior_struct(prior)
        rc = self.lib.ref_render_prior(
            N, _c(scene["means"], np.float32), _c(scene["scales"], np.float32), _c(scene["quats"], np.float32),
            _c(scene["dc"], np.float32), _c(scene["rest"], np.float32), rb, _c(scene["opac"], np.float32),
            _c(cam["viewmat"], np.float32), _c(cam["projmat"], np.float32), cam["fx"], cam["fy"], cam["cx"],
            cam["cy"], H, W, 0.01, scene["degree"], cam["dtu"], _c(cam["cam_pos"], np.float32),
            _c(bg, np.float32), C.byref(st), img, depth.ctypes.data, fT.ctypes.data)
        if rc:
            raise RuntimeError("reference render with a prior failed")
        return img, depth, fT

    def densify(self, bufs, N, cap, gn, vc, ms, half_max_dim, fr_stride, seed, grad_thr=2e-4, size_thr=0.01,
                screen_thr=0.05, check_screen=1, cull_alpha=0.1, cull_scale=0.5, cull_screen=0.15, check_huge=0):
        return self.lib.ref_densify(N, cap, grad_thr, size_thr, screen_thr, check_screen, cull_alpha, cull_scale,
                                    cull_screen, check_huge, _c(gn, np.float32), _c(vc, np.float32),
                                    _c(ms, np.float32), half_max_dim, _ptrs(bufs), fr_stride, seed)

    def opacity_reset(self, o, m1, m2, reset_logit):
        self.lib.ref_opacity_reset(o, m1, m2, o.size, reset_logit)

    def radix_sort(self, keys, vals, key_bits):
        k = _c(keys, np.uint64).copy()
        v = _c(vals, np.int32).copy()
        self.lib.ref_radix_sort(k, v, k.size, key_bits)
        return k, v

    def knn3(self, pts):
        pts = _c(pts, np.float32)
        out = np.zeros(pts.shape[0], np.float32)
        self.lib.ref_knn3(pts, pts.shape[0], out)
        return out

    def init_gaussians(self, means, rgb, seed=42, logit=None):
        n = means.shape[0]
        logit = np.float32(np.log(np.float32(0.1) / np.float32(0.9))) if logit is None else logit
        s, q = np.zeros((n, 3), np.float32), np.zeros((n, 4), np.float32)
        d, o = np.zeros((n, 3), np.float32), np.zeros((n, 1), np.float32)
        self.lib.ref_init_gaussians(_c(means, np.float32), _c(rgb, np.uint8), n, seed, logit, s, q, d, o)
        return s, q, d, o

    def image_metrics(self, a, b):
        out = np.zeros(3, np.float64)
        self.lib.ref_image_metrics(_c(a, np.float32), _c(b, np.float32), a.shape[0], a.shape[1], out)
        return out

    def pack_rgba8(self, img):
        n = img.shape[0] * img.shape[1]
        out = np.zeros((img.shape[0], img.shape[1], 4), np.uint8)
        self.lib.ref_pack_rgba8(_c(img, np.float32), n, out)
        return out


def _ref_resize_area(self, src, dw, dh):
    src = np.ascontiguousarray(src)
    out = np.zeros((dh, dw, 3), np.float32)
    self.lib.ref_resize_area(src.ctypes.data, int(src.dtype == np.uint8), src.shape[1], src.shape[0], out, dw, dh)
    return out


def _ref_undistort(self, src, intr, dist, roi_x, roi_y, dw, dh):
    out = np.zeros((dh, dw, 3), np.float32)
    self.lib.ref_undistort(_c(src, np.float32), src.shape[1], src.shape[0], _c(intr, np.float32),
                          _c(dist, np.float32), roi_x, roi_y, out, dw, dh)
    return out


def _ref_save(self, scene, path, step=0, scale=1.0, translation=(0.0, 0.0, 0.0), keep_crs=False, splat=False):
    """saveGaussianPly (or saveGaussianSplat) of the reference on a host scene."""
    n = scene["means"].shape[0]
    rb = sh_bases(scene["degree"]) - 1
    arrs = [_c(scene[k], np.float32) for k in ("means", "scales", "quats", "dc")]
    arrs.append(_c(np.asarray(scene["rest"], np.float32).reshape(n, rb, 3), np.float32))
    arrs.append(_c(scene["opac"], np.float32))
    self.lib.ref_save_ply.argtypes = [C.c_int, vp, C.c_int, C.c_char_p, C.c_int, C.c_float, f32p, C.c_int, C.c_int]
    if self.lib.ref_save_ply(n, _ptrs(arrs), rb, path.encode(), step, scale, np.asarray(translation, np.float32),
                             int(keep_crs), int(splat)):
        raise RuntimeError("reference save failed")


def _ref_load_ply(self, path, n, rest_bases, scale=1.0, translation=(0.0, 0.0, 0.0), keep_crs=False):
    out = dict(means=np.zeros((n, 3), np.float32), scales=np.zeros((n, 3), np.float32),
               quats=np.zeros((n, 4), np.float32), dc=np.zeros((n, 3), np.float32),
               rest=np.zeros((n, rest_bases, 3), np.float32), opac=np.zeros((n, 1), np.float32))
    step = C.c_int(0)
    self.lib.ref_load_ply.argtypes = [C.c_char_p, C.c_float, f32p, C.c_int, C.c_int, C.c_int, vp, vp]
    rc = self.lib.ref_load_ply(path.encode(), scale, np.asarray(translation, np.float32), int(keep_crs), n, rest_bases,
                               _ptrs([out[k] for k in ("means", "scales", "quats", "dc", "rest", "opac")]),
                               C.byref(step))
    if rc:
        raise RuntimeError("reference load failed")
    out["step"] = step.value
    return out


Reference.save = _ref_save
Reference.load_ply = _ref_load_ply
Reference.resize_area = _ref_resize_area
Reference.undistort = _ref_undistort


def host_isa() -> str:
    """'v4' when this host runs x86-64-v4 code (AVX-512 F/BW/CD/DQ/VL), else 'v3'."""
    try:
        flags = next((l for l in open("/proc/cpuinfo") if l.startswith("flags")), "").split()
    except OSError:
        return "v3"
    return "v4" if all(f in flags for f in ("avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl")) else "v3"


def reference_available(perf: bool = False, isa: str = "v3") -> bool:
    name = "libmsplat_ref.so" if not perf else ("libmsplat_ref_perf_v4.so" if isa == "v4" else
                                                "libmsplat_ref_perf.so")
    return os.path.exists(os.path.join(_HERE, "_ref", name))
