"""A smooth 64-Gaussian scene on a 32x32 image and central finite differences (h = 1e-3)
of the FP64 CPU renderer, for gradient checks of every parameter.

Smooth as in the reference's tests/gradcheck.cpp: every Gaussian spans the whole image
(2D std 20-35 px, opacity 0.05-0.1), so alpha stays above 1/255 on every pixel (6x
margin), T stays above 1e-4, the tile rectangles cover the whole image, and the depths
are 0.025 apart, so no shift of h reorders the sort or crosses a cutoff. Two losses: a
linear functional <image, R> and the training loss w (1 - SSIM) + (1 - w) L1, both in
float64. The SH view direction is not differentiated (core/cpu/msplat_cpu.cpp:1111-1137),
so means are only checked when the SH rest coefficients are zero.
"""
import numpy as np

GAUSS = np.array([0.0010283801, 0.0075987581, 0.0360007721, 0.1093606895, 0.2130055377, 0.2660117249,
                  0.2130055377, 0.1093606895, 0.0360007721, 0.0075987581, 0.0010283801])
BG = np.array([0.2, 0.3, 0.45])
EYE = (0.4, -0.3, 4.0)
KEYS = ("means", "scales", "quats", "dc", "rest", "opac")
SSIM_WEIGHT = 0.8
H_FD = 1e-3


def scene(view_dependent: bool, seed=1234):
    rng = np.random.default_rng(seed)
    n = 64
    eye, fwd = np.array(EYE), -np.array(EYE) / np.linalg.norm(EYE)
    right = np.cross(fwd, (0.0, 1.0, 0.0))
    right /= np.linalg.norm(right)
    up = np.cross(right, fwd)
    d = 3.2 + 0.025 * rng.permutation(n)  # distinct depths in shuffled order
    u, v = rng.uniform(-0.1, 0.1, n) * d, rng.uniform(-0.1, 0.1, n) * d
    means = eye + d[:, None] * fwd + u[:, None] * right + v[:, None] * up
    q = rng.uniform(-1, 1, (n, 4))
    q = 1.3 * q / np.linalg.norm(q, axis=1, keepdims=True)  # not unit: the normalisation is differentiated
    return dict(
        means=means.astype(np.float32),
        scales=np.log(rng.uniform(3.0, 4.0, (n, 3))).astype(np.float32),
        quats=q.astype(np.float32),
        dc=rng.uniform(-0.6, 0.6, (n, 3)).astype(np.float32),
        rest=(rng.uniform(-0.15, 0.15, (n, 3, 3)) if view_dependent else np.zeros((n, 3, 3))).astype(np.float32),
        opac=rng.uniform(-3.0, -2.2, (n, 1)).astype(np.float32),
        degree=1)


def camera():
    from msplat.camera import look_at, make_camera
    return make_camera(look_at(EYE, (0.0, 0.0, 0.0)), 32, 32, 30.0, 29.0, 16.3, 15.6, dtu=1)


def linear_weights():
    return np.random.default_rng(7).normal(0, 1, (32, 32, 3))


def target(oracle, sc, cam):
    """The FP64 render plus a +-(0.02..0.08) pattern: no pixel sits on the L1 kink."""
    base = oracle.render(sc, cam, BG, dtype=np.float64)[2]["img"]
    yy, xx = np.mgrid[0:32, 0:32]
    off = np.stack([np.where((xx // 3 + yy // 5 + c) % 2, 1.0, -1.0) *
                    (0.05 + 0.03 * np.sin(0.3 * xx + c) * np.cos(0.2 * yy)) for c in range(3)], -1)
    return base + off


def _blur(a):
    """zero-padded separable 11-tap blur of (C, H, W)"""
    H, W = a.shape[-2:]
    p = np.pad(a, ((0, 0), (0, 0), (5, 5)))
    t = sum(GAUSS[k] * p[..., k:k + W] for k in range(11))
    p = np.pad(t, ((0, 0), (5, 5), (0, 0)))
    return sum(GAUSS[k] * p[:, k:k + H, :] for k in range(11))


def train_loss(img, gt, w=SSIM_WEIGHT):
    """mean over pixels and channels of w (1 - SSIM) + (1 - w) |gt - img| (zero-padded SSIM)"""
    x, y = np.moveaxis(gt, -1, 0).astype(np.float64), np.moveaxis(img, -1, 0).astype(np.float64)
    mx, my = _blur(x), _blur(y)
    vx, vy, cxy = _blur(x * x) - mx * mx, _blur(y * y) - my * my, _blur(x * y) - mx * my
    C1, C2 = 1e-4, 9e-4
    ssim = ((2 * mx * my + C1) * (2 * cxy + C2)) / ((mx * mx + my * my + C1) * (vx + vy + C2))
    return w * (1 - ssim.mean()) + (1 - w) * np.abs(x - y).mean()


def finite_differences(oracle, sc, cam, loss):
    """{key: dL/dparam} by central differences of the FP64 renderer, every element."""
    R, gt = linear_weights(), target(oracle, sc, cam)

    def f(s):
        img = oracle.render(s, cam, BG, dtype=np.float64)[2]["img"]
        return float((img * R).sum()) if loss == "linear" else float(train_loss(img, gt))

    s64 = {k: (v.astype(np.float64) if isinstance(v, np.ndarray) else v) for k, v in sc.items()}
    out = {}
    for k in KEYS:
        if k == "means" and sc["rest"].any():
            continue
        fd = np.zeros_like(s64[k])
        flat = s64[k].reshape(-1)
        for i in range(flat.size):
            v = flat[i]
            flat[i] = v + H_FD
            lp = f(s64)
            flat[i] = v - H_FD
            lm = f(s64)
            flat[i] = v
            fd.reshape(-1)[i] = (lp - lm) / (2 * H_FD)
        out[k] = fd
    return out


def compare(analytic, fds, label, rel_tol=1e-3):
    """Per group: relative L2 error < rel_tol, and every element within rel_tol of its
    finite difference plus 1e-4 of the group's largest (the FD truncation error of
    elements near zero)."""
    checked = 0
    for k, fd in fds.items():
        a = np.asarray(analytic[k], np.float64).reshape(fd.shape)
        scale = np.abs(fd).max()
        err = np.abs(a - fd)
        bad = err > rel_tol * np.abs(fd) + 1e-4 * scale
        rel = np.linalg.norm(a - fd) / max(np.linalg.norm(fd), 1e-30)
        print(f"{label} {k}: relative L2 {rel:.2e}, worst {err.max() / max(scale, 1e-30):.2e} of max, "
              f"{bad.sum()} / {fd.size} outside")
        assert rel < rel_tol, (label, k, rel)
        assert bad.sum() == 0, (label, k, np.argwhere(bad)[:5])
        checked += fd.size
    return checked
