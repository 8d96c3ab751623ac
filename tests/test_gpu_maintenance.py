"""Densify/prune, opacity reset, kNN scale initialisation, image metrics and RGBA8
packing on the GPU (through the C ABI) against the oracle, the golden vectors of the
compiled reference, and the reference's own brute-force and numpy checks
(tests/test_msplat.py:444-597).

Bars: gaussian counts, split/duplicate/keep decisions (through the compacted order) and
every integer or copied value bit-exact; split-child offsets within 2e-6 relative (the
Box-Muller normals go through log/sin/cos, rounded once from double on the GPU, glibc's
float routines on the CPU); kNN distances bit-exact; metrics bit-exact except SSIM/PSNR
through log10 (double, 1e-12).
"""
import ctypes as C

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu


def _run_densify(ctx, bufs_np, N, cap, gn, vc, ms, half, fr, seed, **kw):
    from msplat._abi import DeviceArray, check, lib
    dev = [DeviceArray.from_numpy(ctx, b) for b in bufs_np]
    dgn, dvc, dms = (DeviceArray.from_numpy(ctx, np.ascontiguousarray(x, np.float32)) for x in (gn, vc, ms))
    arr = (C.c_void_p * 18)(*[d.ptr for d in dev])
    out = C.c_int32()
    p = dict(grad_thr=2e-4, size_thr=0.01, screen_thr=0.05, check_screen=1, cull_alpha=0.1, cull_scale=0.5,
             cull_screen=0.15, check_huge=0)
    p.update(kw)
    check(lib().msplat_hip_densify(ctx.ptr, N, cap, p["grad_thr"], p["size_thr"], p["screen_thr"], p["check_screen"],
                                   p["cull_alpha"], p["cull_scale"], p["cull_screen"], p["check_huge"], dgn.ptr,
                                   dvc.ptr, dms.ptr, half, arr, fr, seed, C.byref(out)))
    return out.value, [d.numpy() for d in dev]


def _compare_densified(got, ref, n, N_children_from):
    for i in range(18):
        a, b = got[i][:n], ref[i][:n]
        if i == 0:  # means: split children carry the RNG offsets
            np.testing.assert_allclose(a, b, rtol=2e-6, atol=1e-7, err_msg="means")
        else:
            np.testing.assert_array_equal(a, b, err_msg=f"buffer {i}")


@pytest.mark.parametrize("huge", [0, 1])
def test_densify_matches_golden(ctx, huge):
    d = golden("densify")
    N, fr = int(d["N"]), int(d["fr"])
    cap = 3 * N
    bufs = []
    for i in range(18):
        b = np.zeros((cap, d[f"in{i}"].shape[1]), np.float32)
        b[:N] = d[f"in{i}"]
        bufs.append(b)
    n, got = _run_densify(ctx, bufs, N, cap, d["gn"], d["vc"], d["ms"], 40.0, fr, int(d["seed"]), check_huge=huge)
    assert n == int(d[f"count_huge{huge}"])
    ref = [d[f"out{huge}_{i}"] for i in range(18)]
    _compare_densified(got, ref, n, N)


def test_densify_matches_oracle_large(ctx, oracle):
    rng = np.random.default_rng(5)
    N, fr = 20000, 45
    cap = 3 * N + 11
    widths = [3, 3, 4, 3, fr, 1]
    bufs = []
    for i in range(18):
        b = np.zeros((cap, widths[i % 6]), np.float32)
        b[:N] = rng.normal(0, 1, (N, widths[i % 6]))
        if i == 1:
            b[:N] = rng.normal(-4, 1.2, (N, 3))
        if i >= 6:
            b[:N] = np.abs(b[:N]) * 0.01
        bufs.append(b)
    vc = rng.integers(0, 6, N).astype(np.float32)
    gn = (rng.random(N) * 4e-3 * vc).astype(np.float32)
    ms = (rng.random(N) * 0.25).astype(np.float32)
    ref_bufs = [b.copy() for b in bufs]
    n_ref, split, dup, keep = oracle.densify(ref_bufs, N, cap, gn, vc, ms, 632.0, fr, 1234567, cull_alpha=0.005,
                                             check_huge=1)
    n, got = _run_densify(ctx, bufs, N, cap, gn, vc, ms, 632.0, fr, 1234567, cull_alpha=0.005, check_huge=1)
    assert split.sum() > 100 and dup.sum() > 100 and (keep == 0).sum() > 100
    assert n == n_ref
    _compare_densified(got, ref_bufs, n, N)


def test_opacity_reset(ctx, oracle):
    from msplat._abi import DeviceArray, check, lib
    rng = np.random.default_rng(2)
    o, m1, m2 = (rng.normal(0, 2, 1000).astype(np.float32) for _ in range(3))
    d = [DeviceArray.from_numpy(ctx, x) for x in (o, m1, m2)]
    check(lib().msplat_hip_opacity_reset(ctx.ptr, d[0].ptr, d[1].ptr, d[2].ptr, 1000, -1.3862943611198906))
    oracle.opacity_reset(o, m1, m2, np.float32(-1.3862943611198906))
    for a, b in zip(d, (o, m1, m2)):
        np.testing.assert_array_equal(a.numpy(), b)


def test_opacity_reset_matches_golden(ctx):
    """HIP against the compiled reference's own outputs (ties with the cap, +-inf, -0)."""
    from conftest import golden
    from msplat._abi import DeviceArray, check, lib
    g = golden("opacity_reset")
    d = [DeviceArray.from_numpy(ctx, g[k]) for k in ("o", "m1", "m2")]
    check(lib().msplat_hip_opacity_reset(ctx.ptr, d[0].ptr, d[1].ptr, d[2].ptr, g["o"].size, float(g["reset"])))
    for a, k in zip(d, ("out_o", "out_m1", "out_m2")):
        np.testing.assert_array_equal(a.numpy(), g[k])


def _knn_points(kind, rng):
    if kind == "uniform":
        return rng.uniform(-1, 1, (3000, 3))
    if kind == "street":  # thin road plane + facade + sparse far points
        road = np.c_[rng.uniform(-20, 20, 2000), rng.normal(0, 0.02, 2000), rng.uniform(0, 80, 2000)]
        facade = np.c_[8 + rng.normal(0, 0.05, 800), rng.uniform(0, 10, 800), rng.uniform(0, 80, 800)]
        return np.vstack([road, facade, rng.normal(0, 300, (200, 3))])
    return np.vstack([rng.uniform(-1, 1, (500, 3)), np.repeat(rng.uniform(-1, 1, (8, 3)), 4, axis=0)])


@pytest.mark.parametrize("kind", ["uniform", "street", "duplicates"])
def test_knn_matches_bruteforce_and_oracle(ctx, oracle, kind):
    from msplat._abi import DeviceArray, check, lib
    pts = _knn_points(kind, np.random.default_rng(7)).astype(np.float32)
    dp = DeviceArray.from_numpy(ctx, pts)
    out = DeviceArray(ctx, (pts.shape[0],))
    check(lib().msplat_hip_knn3_mean_dist(ctx.ptr, dp.ptr, pts.shape[0], out.ptr))
    got = out.numpy()
    d2 = ((pts[:, None, :].astype(np.float64) - pts[None, :, :]) ** 2).sum(-1)
    np.fill_diagonal(d2, np.inf)
    np.testing.assert_allclose(got, np.sqrt(np.sort(d2, axis=1)[:, :3]).mean(axis=1), rtol=1e-4, atol=1e-6)
    np.testing.assert_array_equal(got, oracle.knn3(pts))


@pytest.mark.parametrize("n", [1, 2, 3, 4, 6])
def test_knn_tiny_clouds(ctx, oracle, n):
    """Fewer than three neighbours: the mean runs over those there are, and a lone point
    gets 1.0 (core/cpu/msplat_cpu.cpp:1572-1576). n = 6 repeats one point: all zero."""
    from msplat._abi import DeviceArray, check, lib
    rng = np.random.default_rng(n)
    pts = rng.uniform(-1, 1, (n, 3)).astype(np.float32) if n < 6 else np.repeat(
        rng.uniform(-1, 1, (1, 3)).astype(np.float32), n, axis=0)
    dp = DeviceArray.from_numpy(ctx, pts)
    out = DeviceArray(ctx, (n,))
    check(lib().msplat_hip_knn3_mean_dist(ctx.ptr, dp.ptr, n, out.ptr))
    np.testing.assert_array_equal(out.numpy(), oracle.knn3(pts))
    if n == 1:
        assert out.numpy()[0] == 1.0
    if n == 6:
        assert (out.numpy() == 0.0).all()


def test_knn_large_street_cloud(ctx, oracle):
    """The street recipe at 200k points: bit-exact against the oracle's grid search."""
    from msplat._abi import DeviceArray, check, lib
    from msplat.scenes import street_points
    pts, _ = street_points(3, 200_000)
    dp = DeviceArray.from_numpy(ctx, pts)
    out = DeviceArray(ctx, (pts.shape[0],))
    check(lib().msplat_hip_knn3_mean_dist(ctx.ptr, dp.ptr, pts.shape[0], out.ptr))
    np.testing.assert_array_equal(out.numpy(), oracle.knn3(pts))


def test_init_gaussians_matches_golden(ctx):
    from msplat._abi import DeviceArray, check, lib
    d = golden("knn_init")
    n = d["points"].shape[0]
    dp, dc_rgb = DeviceArray.from_numpy(ctx, d["points"]), DeviceArray.from_numpy(ctx, d["rgb"])
    s, q, dc, o = DeviceArray(ctx, (n, 3)), DeviceArray(ctx, (n, 4)), DeviceArray(ctx, (n, 3)), DeviceArray(ctx, (n, 1))
    logit = float(np.float32(np.log(np.float32(0.1) / np.float32(0.9))))
    check(lib().msplat_hip_init_gaussians(ctx.ptr, dp.ptr, dc_rgb.ptr, n, 42, logit, s.ptr, q.ptr, dc.ptr, o.ptr))
    # log/sin/cos rounded once from double on the GPU, glibc's float routines in the reference
    np.testing.assert_array_max_ulp(s.numpy(), d["scales"], maxulp=1)
    np.testing.assert_allclose(q.numpy(), d["quats"], rtol=0, atol=2e-7)
    np.testing.assert_array_equal(dc.numpy(), d["dc"])
    np.testing.assert_array_equal(o.numpy(), d["opac"])


@pytest.mark.parametrize("shape", [(37, 53), (5, 40), (64, 64), (100, 17)])
def test_image_metrics(ctx, oracle, shape):
    from msplat._abi import DeviceArray, check, lib
    rng = np.random.default_rng(shape[0] + shape[1])
    a = rng.random((*shape, 3), dtype=np.float32)
    b = np.clip(a + rng.normal(0, 0.05, a.shape), 0, 1).astype(np.float32)
    da, db = DeviceArray.from_numpy(ctx, a), DeviceArray.from_numpy(ctx, b)
    out = (C.c_double * 3)()
    check(lib().msplat_hip_image_metrics(ctx.ptr, da.ptr, db.ptr, shape[0], shape[1], out))
    ref = oracle.image_metrics(a, b)
    assert out[1] == ref[1] and out[2] == ref[2]
    assert abs(out[0] - ref[0]) <= 1e-12 * abs(ref[0])


def test_image_metrics_golden(ctx):
    from msplat._abi import DeviceArray, check, lib
    d = golden("image_ops")
    da, db = DeviceArray.from_numpy(ctx, d["a"]), DeviceArray.from_numpy(ctx, d["b"])
    out = (C.c_double * 3)()
    check(lib().msplat_hip_image_metrics(ctx.ptr, da.ptr, db.ptr, 37, 53, out))
    np.testing.assert_allclose(list(out), d["metrics"], rtol=1e-12)


def test_pack_rgba8_matches_golden(ctx):
    from msplat._abi import DeviceArray, check, lib
    d = golden("image_ops")
    img = d["img"]
    di = DeviceArray.from_numpy(ctx, img)
    out = DeviceArray(ctx, (img.shape[0], img.shape[1], 4), np.uint8)
    check(lib().msplat_hip_pack_rgba8(ctx.ptr, di.ptr, img.shape[0] * img.shape[1], out.ptr))
    np.testing.assert_array_equal(out.numpy(), d["rgba"])
