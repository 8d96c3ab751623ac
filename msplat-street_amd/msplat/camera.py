"""Camera setup for one view: view/projection matrices, SH degree in use, tile grid.

Follows Model::prepareCam (core/src/model.cpp:498-546) in float32: the camera-to-world
pose is in the OpenGL convention (columns right, up, back), flipped to OpenCV (y down,
z forward) before inversion; the projection uses znear 0.001 and zfar 1000 with the
principal point applied later in pixel space. The field of view goes through the C
library's atanf/tanf, the functions the reference's std::atan/std::tan on floats call,
so the matrices equal the reference's bit for bit. At a downscale factor the image
size is the pyramid level's (levelSize, image_sizing.hpp:62) and the intrinsics scale
by the per-axis size ratio (model.cpp:502-508).
"""
from __future__ import annotations

import ctypes as C
import ctypes.util

import numpy as np

TILE = 16
F = np.float32

_libm = C.CDLL(ctypes.util.find_library("m"))
for _name in ("atanf", "tanf", "expf", "logf"):
    getattr(_libm, _name).argtypes = [C.c_float]
    getattr(_libm, _name).restype = C.c_float


def atanf(x) -> np.float32:
    return F(_libm.atanf(float(F(x))))


def tanf(x) -> np.float32:
    return F(_libm.tanf(float(F(x))))


def expf(x) -> np.float32:
    return F(_libm.expf(float(F(x))))


def logf(x) -> np.float32:
    return F(_libm.logf(float(F(x))))


def level_size(full: int, factor: int) -> int:
    """levelSize (image_sizing.hpp:62): the pyramid level's whole-pixel size."""
    return int(full) if factor <= 1 else int(full) // int(factor)


def look_at(eye, target, up=(0.0, 1.0, 0.0)) -> np.ndarray:
    """OpenGL camera-to-world (4x4 row-major) at `eye` looking at `target`."""
    eye, target, up = (np.asarray(v, np.float64) for v in (eye, target, up))
    f = target - eye
    f /= np.linalg.norm(f)
    r = np.cross(f, up)
    r /= np.linalg.norm(r)
    u = np.cross(r, f)
    m = np.eye(4)
    m[:3, 0], m[:3, 1], m[:3, 2], m[:3, 3] = r, u, -f, eye
    return m.astype(np.float32)


def schedule_downscale(step: int, num_downscales: int, resolution_schedule: int) -> int:
    """scheduleDownscaleFactor (core/src/image_sizing.cpp:129-133)."""
    if num_downscales <= 0 or resolution_schedule <= 0:
        return 1
    remaining = num_downscales - step // resolution_schedule
    return 1 << min(max(remaining, 0), 30)


def view_matrices(cam_to_world, W, H, fx, fy):
    """(viewmat, projmat, cam_pos) as prepareCam builds them (model.cpp:509-525), all
    float32: every product and sum rounds to float32, left to right."""
    d = np.asarray(cam_to_world, F).reshape(16)
    R = np.zeros((3, 3), F)
    T = np.zeros(3, F)
    for i in range(3):
        R[i, 0], R[i, 1], R[i, 2], T[i] = d[i * 4], -d[i * 4 + 1], -d[i * 4 + 2], d[i * 4 + 3]
    Rinv = R.T.copy()
    Tinv = np.array([-(Rinv[i, 0] * T[0] + Rinv[i, 1] * T[1] + Rinv[i, 2] * T[2]) for i in range(3)], F)
    vm = np.array([Rinv[0, 0], Rinv[0, 1], Rinv[0, 2], Tinv[0], Rinv[1, 0], Rinv[1, 1], Rinv[1, 2], Tinv[1],
                   Rinv[2, 0], Rinv[2, 1], Rinv[2, 2], Tinv[2], 0, 0, 0, 1], F)
    fovx = F(2.0) * atanf(F(W) / (F(2.0) * F(fx)))
    fovy = F(2.0) * atanf(F(H) / (F(2.0) * F(fy)))
    t_p = F(0.001) * tanf(F(0.5) * fovy)
    r_p = F(0.001) * tanf(F(0.5) * fovx)
    zn, zf = F(0.001), F(1000.0)
    pm = np.array([zn / r_p, 0, 0, 0, 0, zn / t_p, 0, 0, 0, 0, (zf + zn) / (zf - zn), -zf * zn / (zf - zn),
                   0, 0, 1, 0], F)
    pvm = np.zeros(16, F)
    for i in range(4):
        for j in range(4):
            acc = F(0)
            for k in range(4):
                acc = F(acc + pm[i * 4 + k] * vm[k * 4 + j])
            pvm[i * 4 + j] = acc
    return vm, pvm, T.copy()


def make_camera(cam_to_world, W, H, fx, fy, cx=None, cy=None, dtu=0, downscale=1):
    """Camera dict in the layout every backend entry point takes. With downscale > 1 the
    view is the pyramid level prepareCam renders at that factor (model.cpp:500-508)."""
    cx = W / 2.0 if cx is None else cx
    cy = H / 2.0 if cy is None else cy
    if downscale > 1:
        w, h = level_size(W, downscale), level_size(H, downscale)
        rx = F(w) / F(W) if W > 0 else F(1) / F(downscale)
        ry = F(h) / F(H) if H > 0 else F(1) / F(downscale)
        fx, cx, fy, cy = F(fx) * rx, F(cx) * rx, F(fy) * ry, F(cy) * ry
        W, H = w, h
    vm, pm, pos = view_matrices(cam_to_world, W, H, fx, fy)
    return dict(viewmat=vm, projmat=pm, cam_pos=pos, fx=F(fx), fy=F(fy), cx=F(cx), cy=F(cy), W=int(W), H=int(H),
                tiles_x=(int(W) + TILE - 1) // TILE, tiles_y=(int(H) + TILE - 1) // TILE, dtu=int(dtu),
                cam_to_world=np.asarray(cam_to_world, F))


def degrees_to_use(step: int, sh_degree: int, sh_degree_interval: int) -> int:
    """degreesToUse = min(step / shDegreeInterval, shDegree) (model.cpp:535)."""
    return min(step // max(sh_degree_interval, 1), sh_degree)
