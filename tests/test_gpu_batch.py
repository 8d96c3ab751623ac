"""Batched views on the GPU: one projection, sort, persistent (view, tile) rasterizer,
loss and backward over several views of mixed resolution, against the one-view entry
points (which the oracle tests pin): images bit for bit, gradients within float
rounding of the per-view sum, the one-view batch bit for bit equal to the one-view
train step, graph replays bit for bit equal to eager steps, and the capacity regrowth."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BG = (0.2, 0.35, 0.5)
SIZES = [(64, 48), (97, 71), (128, 128), (50, 70), (33, 17), (160, 90)]  # (W, H), not all tile multiples


@pytest.fixture(autouse=True)
def _per_tile_one_view_frames(ctx):
    """The batched rasterizer runs one workgroup per (view, tile); the one-view frames it
    is compared with bit for bit run the same per-tile form (no depth chunks, which give
    the same frame up to float rounding: tests/test_gpu_chunked.py)."""
    from msplat._abi import TUNE_CHUNKS
    ctx.set_tuning(TUNE_CHUNKS, 1)
    yield
    ctx.set_tuning(TUNE_CHUNKS, 0)


def _setup(seed=3, n=4000, deg=3, sizes=SIZES, dtu=3):
    from msplat.camera import look_at, make_camera
    from msplat.scenes import random_scene
    sc = random_scene(seed, n, degree=deg, log_scale=(-3.2, 0.4))
    rng = np.random.default_rng(seed + 1)
    cams = []
    for k, (W, H) in enumerate(sizes):
        th = 2 * np.pi * k / len(sizes) + rng.uniform(-0.2, 0.2)
        eye = (3.2 * np.cos(th), rng.uniform(-0.5, 0.5), 3.2 * np.sin(th))
        f = 0.9 * max(W, H)
        cams.append(make_camera(look_at(eye, (0, 0, 0)), W, H, f, f * 1.02, W / 2 + 0.3, H / 2 - 0.2, dtu=dtu))
    return sc, cams


def _close(a, b, rtol):
    scale = max(np.abs(b).max(), 1e-30)
    err = np.abs(a - b).max() / scale
    assert err <= rtol, err
    return err


@pytest.mark.parametrize("deg,dtus", [(3, (3,)), (3, (0, 1, 3, 2, 1, 0)), (4, (4, 2, 0, 4, 1, 3)), (2, (0,))])
def test_render_batch_equals_single_views(ctx, deg, dtus):
    """Per-view images bit for bit, also when the views use different SH bands (the batch
    stages the most any view uses; degree 4 reads in place)."""
    from msplat.backend import BatchViews, DeviceGaussians, render_batch_numpy, render_numpy
    sc, cams = _setup(deg=deg)
    for k, cam in enumerate(cams):
        cam["dtu"] = dtus[k % len(dtus)]
    dg = DeviceGaussians(ctx, sc)
    bv = BatchViews(cams, deg, [BG])
    imgs = render_batch_numpy(ctx, dg, bv)
    for cam, img in zip(cams, imgs):
        ref = render_numpy(ctx, dg, cam, BG)
        np.testing.assert_array_equal(img, ref)
        assert img.std() > 0.01  # something was drawn


def test_render_backward_batch_is_the_sum_of_views(ctx):
    from msplat._abi import DeviceArray
    from msplat.backend import BatchViews, DeviceGaussians, render, render_backward, render_backward_batch, \
        render_batch
    sc, cams = _setup(seed=5)
    dg = DeviceGaussians(ctx, sc)
    bv = BatchViews(cams, 3, [BG])
    rng = np.random.default_rng(6)
    v_imgs = [(rng.normal(0, 1, (c["H"], c["W"], 3)) * 1e-3).astype(np.float32) for c in cams]
    total = {}
    for cam, v in zip(cams, v_imgs):
        render(ctx, dg, cam, BG)
        gr = render_backward(ctx, dg, cam, BG, DeviceArray.from_numpy(ctx, v))
        for k, x in gr.items():
            total[k] = total.get(k, 0) + x.astype(np.float64)
    render_batch(ctx, dg, bv)
    packed = DeviceArray.from_numpy(ctx, np.concatenate([v.reshape(-1, 3) for v in v_imgs]))
    got = render_backward_batch(ctx, dg, bv, packed)
    for k in total:
        err = _close(got[k].astype(np.float64), total[k], 2e-5)
        print(f"{k}: max error {err:.2e} of the largest")



def test_repeated_backward_batch_of_one_frame(ctx):
    """Two backward passes of one batched render, with cotangents zero in different
    views, give the gradients of a fresh render and backward each (the second pass
    re-flags the slots the first one wrote)."""
    from msplat._abi import DeviceArray
    from msplat.backend import BatchViews, DeviceGaussians, render_backward_batch, render_batch
    sc, cams = _setup(seed=7)
    dg = DeviceGaussians(ctx, sc)
    bv = BatchViews(cams, 3, [BG])
    rng = np.random.default_rng(8)
    v = [(rng.normal(0, 1, (c["H"], c["W"], 3)) * 1e-3).astype(np.float32) for c in cams]
    half = len(cams) // 2
    va = np.concatenate([(x if i < half else 0 * x).reshape(-1, 3) for i, x in enumerate(v)])
    vb = np.concatenate([(0 * x if i < half else x).reshape(-1, 3) for i, x in enumerate(v)])
    da, db = DeviceArray.from_numpy(ctx, va), DeviceArray.from_numpy(ctx, vb)
    fresh = []
    for d in (da, db):
        render_batch(ctx, dg, bv)
        fresh.append(render_backward_batch(ctx, dg, bv, d))
    render_batch(ctx, dg, bv)
    again = [render_backward_batch(ctx, dg, bv, d) for d in (da, db, da)]
    for got, want in zip(again, fresh + fresh[:1]):
        for k in want:
            assert np.abs(want[k]).max() > 0, k
            np.testing.assert_array_equal(got[k], want[k], err_msg=k)

def _state(ctx, sc):
    from msplat.backend import DeviceGaussians
    return DeviceGaussians(ctx, sc)


def _targets(ctx, cams, seed=9):
    from msplat._abi import DeviceArray
    rng = np.random.default_rng(seed)
    return [DeviceArray.from_numpy(ctx, rng.uniform(0, 1, (c["H"], c["W"], 3)).astype(np.float32)) for c in cams]


def test_one_view_batch_equals_train_step(ctx):
    """the batch path on one view is the one-view step, bit for bit"""
    from msplat.backend import BatchViews, adam_coeffs, train_step, train_step_batch
    sc, cams = _setup(seed=7, sizes=[(97, 71)])
    gts = _targets(ctx, cams)
    a, b = _state(ctx, sc), _state(ctx, sc)
    lrs = [1.6e-4, 5e-3, 1e-3, 2.5e-3, 2.5e-3 / 20, 5e-2]
    for t in range(1, 4):
        steps, bc2 = adam_coeffs(lrs, t)
        la = train_step(ctx, a, cams[0], BG, gts[0], steps, bc2, ssim_weight=0.2)
        lb = train_step_batch(ctx, b, BatchViews(cams, 3, [BG], gts), steps, bc2, ssim_weight=0.2)
        assert la == lb[0]
    sa, sb = a.to_numpy(), b.to_numpy()
    for k in sa:
        if k != "degree":
            np.testing.assert_array_equal(sa[k], sb[k], err_msg=k)


def test_batch_gradients_and_statistics_are_view_sums(ctx):
    """beta1 = 0 and zero step sizes leave the parameters and make m1 the gradient: the
    batch's m1 is the sum of the one-view m1s, its statistics the one-view sequence's"""
    from msplat.backend import BatchViews, train_step, train_step_batch
    sc, cams = _setup(seed=11)
    gts = _targets(ctx, cams)
    zero, ones = np.zeros(6, np.float32), np.ones(6, np.float32)
    seq = _state(ctx, sc)
    m1_sum, losses = {}, []
    for cam, gt in zip(cams, gts):
        one = _state(ctx, sc)
        losses.append(train_step(ctx, one, cam, BG, gt, zero, ones, ssim_weight=0.2, beta1=0.0))
        s = one.to_numpy()
        for k in ("means", "scales", "quats", "dc", "rest", "opac"):
            m1_sum[k] = m1_sum.get(k, 0) + s["m1_" + k].astype(np.float64)
        train_step(ctx, seq, cam, BG, gt, zero, ones, ssim_weight=0.2)  # statistics of the sequence
    bat = _state(ctx, sc)
    lb = train_step_batch(ctx, bat, BatchViews(cams, 3, [BG], gts), zero, ones, ssim_weight=0.2, beta1=0.0)
    got = bat.to_numpy()
    np.testing.assert_array_equal(lb, np.array(losses, np.float32))
    for k in m1_sum:
        np.testing.assert_array_equal(got[k], sc[k].reshape(got[k].shape), err_msg=k)  # untouched
        err = _close(got["m1_" + k].astype(np.float64), m1_sum[k], 2e-5)
        print(f"m1 {k}: max error {err:.2e} of the largest")
    ref = seq.to_numpy()
    np.testing.assert_array_equal(got["vis"], ref["vis"])
    np.testing.assert_array_equal(got["ms"], ref["ms"])
    np.testing.assert_allclose(got["gn"], ref["gn"], rtol=2e-6, atol=0)


def test_one_view_m1_reference_uses_beta1_one_minus(ctx):
    """train_step's m1 after one step from zero is (1 - beta1) g: with beta1 = 0.9 the
    batch and the one-view step agree on it too (a check of the beta1 plumbing)"""
    from msplat.backend import BatchViews, train_step, train_step_batch
    sc, cams = _setup(seed=12, sizes=[(64, 48)])
    gts = _targets(ctx, cams)
    zero, ones = np.zeros(6, np.float32), np.ones(6, np.float32)
    a, b = _state(ctx, sc), _state(ctx, sc)
    train_step(ctx, a, cams[0], BG, gts[0], zero, ones)
    train_step_batch(ctx, b, BatchViews(cams, 3, [BG], gts), zero, ones)
    np.testing.assert_array_equal(a.to_numpy()["m1_means"], b.to_numpy()["m1_means"])


def test_batch_with_an_empty_view(ctx):
    """A view that sees no Gaussian (all behind its camera) inside a batch: its image is
    the background, the other views render as alone, and the batch's gradients are the
    sum over the views that see something."""
    from msplat.backend import BatchViews, render_batch_numpy, render_numpy, train_step, train_step_batch
    from msplat.camera import look_at, make_camera
    sc, cams = _setup(seed=13, sizes=[(97, 71), (64, 48)])
    away = make_camera(look_at((0.0, 0.0, 5.0), (0.0, 0.0, 10.0)), 80, 60, 70.0, 70.0, dtu=3)
    views = [cams[0], away, cams[1]]
    dg = _state(ctx, sc)
    imgs = render_batch_numpy(ctx, dg, BatchViews(views, 3, [BG]))
    for cam, img in zip(views, imgs):
        np.testing.assert_array_equal(img, render_numpy(ctx, dg, cam, BG))
    np.testing.assert_array_equal(imgs[1], np.broadcast_to(np.float32(BG), imgs[1].shape))
    gts = _targets(ctx, views, seed=14)
    zero, ones = np.zeros(6, np.float32), np.ones(6, np.float32)
    m1_sum, losses = {}, []
    for cam, gt in zip(views, gts):
        one = _state(ctx, sc)
        losses.append(train_step(ctx, one, cam, BG, gt, zero, ones, ssim_weight=0.2, beta1=0.0))
        st = one.to_numpy()
        for k in ("means", "scales", "quats", "dc", "rest", "opac"):
            m1_sum[k] = m1_sum.get(k, 0) + st["m1_" + k].astype(np.float64)
    bat = _state(ctx, sc)
    lb = train_step_batch(ctx, bat, BatchViews(views, 3, [BG], gts), zero, ones, ssim_weight=0.2, beta1=0.0)
    np.testing.assert_array_equal(lb, np.array(losses, np.float32))
    got = bat.to_numpy()
    for k in m1_sum:
        _close(got["m1_" + k].astype(np.float64), m1_sum[k], 2e-5)


def test_graph_replays_equal_eager_steps(ctx):
    """Captured step replayed with new poses, resolutions (within the buffers) and Adam
    scalars, with eager work and syncs between replays, equals the eager steps bit for bit."""
    from msplat._abi import Context
    from msplat.backend import BatchViews, adam_coeffs, render_numpy, train_step_batch
    sc, cams = _setup(seed=13)
    lrs = [1.6e-4, 5e-3, 1e-3, 2.5e-3, 2.5e-3 / 20, 5e-2]
    eager_ctx, graph_ctx = ctx, Context(0)
    graph_ctx.set_batch_graph(True)
    try:
        e, g = _state(eager_ctx, sc), _state(graph_ctx, sc)
        gts_e, gts_g = _targets(eager_ctx, cams), _targets(graph_ctx, cams)
        order = [list(range(6)), [5, 4, 3, 2, 1, 0], [1, 3, 5, 0, 2, 4], [2, 0, 4, 1, 5, 3]]
        for t, perm in enumerate(order, start=1):
            cs = [cams[i] for i in perm]
            steps, bc2 = adam_coeffs(lrs, t)
            le = train_step_batch(eager_ctx, e, BatchViews(cs, 3, [BG], [gts_e[i] for i in perm]), steps, bc2)
            lg = train_step_batch(graph_ctx, g, BatchViews(cs, 3, [BG], [gts_g[i] for i in perm]), steps, bc2)
            np.testing.assert_array_equal(le, lg)
            render_numpy(graph_ctx, g, cams[0], BG)  # eager work on the graph's context between replays
            se, sg = e.to_numpy(), g.to_numpy()
            for k in se:
                if k != "degree":
                    np.testing.assert_array_equal(se[k], sg[k], err_msg=f"step {t} {k}")
        captures, replays = graph_ctx.batch_graph_stats()
        assert captures == 1 and replays == len(order), (captures, replays)
    finally:
        graph_ctx.close()


def test_capacity_regrowth_gives_the_same_frame(ctx):
    """A graph context sized on a sparse scene meets a dense one: the captured step
    exceeds its intersection capacity, changes nothing, and runs again with grown
    buffers; the result equals an eager step on a fresh context, bit for bit. Eager
    render frames size themselves from a count pass and equal too."""
    from msplat._abi import Context
    from msplat.backend import BatchViews, DeviceGaussians, render_batch_numpy, train_step_batch
    sc_small, cams = _setup(seed=17, n=300)
    sc_big, _ = _setup(seed=17, n=20000)
    g = Context(0)
    g.set_batch_graph(True)
    zero, ones = np.zeros(6, np.float32), np.ones(6, np.float32)
    steps = np.full(6, 1e-3, np.float32)
    try:
        small = _state(g, sc_small)
        train_step_batch(g, small, BatchViews(cams, 3, [BG], _targets(g, cams)), steps, ones)
        big_g = _state(g, sc_big)
        lg = train_step_batch(g, big_g, BatchViews(cams, 3, [BG], _targets(g, cams)), steps, ones)
        captures, replays = g.batch_graph_stats()
        assert captures == 3 and replays == 3  # small; big over capacity; big again
        big_e = _state(ctx, sc_big)
        le = train_step_batch(ctx, big_e, BatchViews(cams, 3, [BG], _targets(ctx, cams)), steps, ones)
        np.testing.assert_array_equal(lg, le)
        se, sg = big_e.to_numpy(), big_g.to_numpy()
        for k in se:
            if k != "degree":
                np.testing.assert_array_equal(sg[k], se[k], err_msg=k)
        got = render_batch_numpy(g, big_g, BatchViews(cams, 3, [BG]))
        ref = render_batch_numpy(ctx, big_e, BatchViews(cams, 3, [BG]))
        for a, b in zip(got, ref):
            np.testing.assert_array_equal(a, b)
    finally:
        g.close()
