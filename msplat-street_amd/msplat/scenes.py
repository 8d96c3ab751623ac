"""Seeded synthetic scenes and camera sets for the benchmark workloads.

Gaussian parameters use the reference layouts (model.cpp:52-64): means (N,3),
log-scales (N,3), quaternions (N,4) stored (w,x,y,z), DC color (N,3), SH rest
(N, bases-1, 3), opacity logits (N,1).
"""
from __future__ import annotations

import numpy as np

from .camera import look_at, make_camera

F = np.float32


def sh_bases(degree: int) -> int:
    return {0: 1, 1: 4, 2: 9, 3: 16}.get(degree, 25)


def uniform_quats(rng, n):
    """Shoemake's uniform unit quaternions."""
    u, v, w = rng.random((3, n))
    tau = 2 * np.pi
    q = np.stack([np.sqrt(1 - u) * np.sin(tau * v), np.sqrt(1 - u) * np.cos(tau * v),
                  np.sqrt(u) * np.sin(tau * w), np.sqrt(u) * np.cos(tau * w)], 1)
    return q.astype(F)


def random_scene(seed, n, degree=1, mean_range=1.0, log_scale=(-4.5, 0.5), dc_sigma=0.5, rest_sigma=0.1,
                 opac=(0.0, 1.0)):
    """The small-scene recipe: means uniform in [-r, r]^3, per-axis log-scales normal,
    uniform rotations, opacity logits normal, SH DC and higher bands normal."""
    rng = np.random.default_rng(seed)
    nb = sh_bases(degree)
    return dict(
        means=rng.uniform(-mean_range, mean_range, (n, 3)).astype(F),
        scales=rng.normal(log_scale[0], log_scale[1], (n, 3)).astype(F),
        quats=uniform_quats(rng, n),
        dc=rng.normal(0, dc_sigma, (n, 3)).astype(F),
        rest=rng.normal(0, rest_sigma, (n, nb - 1, 3)).astype(F) if nb > 1 else np.zeros((n, 1, 3), F),
        opac=rng.normal(opac[0], opac[1], (n, 1)).astype(F),
        degree=degree,
    )


def ring_cameras(count=8, radius=4.0, W=128, H=128, fx=120.0, fy=120.0, dtu=0, height=0.0):
    """Pinhole cameras on a ring around the origin, looking at it, up +Y."""
    cams = []
    for k in range(count):
        th = 2 * np.pi * k / count
        eye = (radius * np.cos(th), height, radius * np.sin(th))
        cams.append(make_camera(look_at(eye, (0, 0, 0)), W, H, fx, fy, dtu=dtu))
    return cams


def street_points(seed, n, centers=((0.0, 0.0, 0.0), (10.0, 0.0, 0.0)), depth=(2.0, 60.0)):
    """Street-scale point cloud: directions uniform on the sphere around each capture
    center, depth log-uniform in [2, 60] m, points split evenly over the centers."""
    rng = np.random.default_rng(seed)
    c = np.asarray(centers, np.float64)
    which = np.arange(n) % len(c)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.exp(rng.uniform(np.log(depth[0]), np.log(depth[1]), n))
    pts = c[which] + d * r[:, None]
    rgb = rng.integers(0, 256, (n, 3), dtype=np.uint8)
    return pts.astype(F), rgb


def street_cameras(count=64, W=1264, H=832, fx=900.0, fy=900.0, start=(-5.0, 0.0, 0.0), end=(15.0, 0.0, 0.0),
                   dtu=0):
    """Views evenly spaced on the straight 20 m path through both capture centers.
    Headings cycle forward, left, backward, right (yaw 0, 90, 180, 270 degrees) with a
    slight downward pitch, so each view sees a different part of the scene."""
    s, e = np.asarray(start, np.float64), np.asarray(end, np.float64)
    cams = []
    for k in range(count):
        p = s + (e - s) * (k + 0.5) / count
        yaw = (k % 4) * np.pi / 2
        fwd = np.array([np.cos(yaw), -0.1, np.sin(yaw)])
        cams.append(make_camera(look_at(p, p + fwd), W, H, fx, fy, dtu=dtu))
    return cams


def gmm_cloud(seed, n, center=(0.0, 0.0, 0.0), clusters=64, sigma=0.5, box=40.0):
    """A Gaussian mixture point cloud: `clusters` centres uniform in a cube of side `box`
    (m) around `center`, every point drawn from a uniformly chosen cluster with isotropic
    deviation `sigma` (m); colours per cluster with a little per-point noise."""
    rng = np.random.default_rng(seed)
    cc = np.asarray(center, np.float64) + rng.uniform(-box / 2, box / 2, (clusters, 3))
    col = rng.integers(30, 226, (clusters, 3))
    lab = rng.integers(0, clusters, n)
    pts = (cc[lab] + rng.normal(0, sigma, (n, 3))).astype(F)
    rgb = np.clip(col[lab] + rng.integers(-25, 26, (n, 3)), 0, 255).astype(np.uint8)
    return pts, rgb


def spline_cameras(count=200, W=1920, H=1080, focal_scale=0.9, seed=0, center=(0.0, 0.0, 0.0),
                   radius=(14.0, 22.0), height=(-3.0, 3.0), knots=8, dtu=0):
    """Poses on a closed Catmull-Rom spline through `knots` seeded control points around
    `center` (angles evenly spaced with jitter, distance U(radius), height U(height)),
    sampled at `count` even parameter steps, each looking at `center` (fx = fy =
    focal_scale W)."""
    rng = np.random.default_rng(seed)
    ang = 2 * np.pi * (np.arange(knots) + rng.uniform(-0.25, 0.25, knots)) / knots
    r = rng.uniform(*radius, knots)
    c = np.asarray(center, np.float64)
    ctrl = c + np.stack([r * np.cos(ang), rng.uniform(*height, knots), r * np.sin(ang)], 1)
    cams = []
    for k in range(count):
        u = knots * k / count
        i, t = int(u), u - int(u)
        p0, p1, p2, p3 = (ctrl[(i + j) % knots] for j in (-1, 0, 1, 2))
        eye = 0.5 * (2 * p1 + (p2 - p0) * t + (2 * p0 - 5 * p1 + 4 * p2 - p3) * t * t
                     + (3 * p1 - p0 - 3 * p2 + p3) * t ** 3)
        f = focal_scale * W
        cams.append(make_camera(look_at(eye, c), W, H, f, f, dtu=dtu))
    return cams
