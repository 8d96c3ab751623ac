"""The captured one-view train step (msplat_hip_set_step_graph): replays of one hipGraph
per step shape give the launch-by-launch steps bit for bit — losses, parameters, both
Adam moments and the densification statistics — including when a replay overflows its
intersection capacity (it changes nothing and is run again), across densification (a new
shape: launch by launch, then a new capture) and with depth-chunked frames."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BG = (0.613, 0.0101, 0.3984)
LRS = (1.6e-4, 5e-3, 1e-3, 2.5e-3, 1.25e-4, 5e-2)


def _workload(ctx, n=20000, seed=1000):
    from msplat._abi import DeviceArray
    from msplat.backend import DeviceGaussians, render_numpy
    from msplat.scenes import random_scene, ring_cameras
    cams = ring_cameras(8, 4.0, 128, 128, 120.0, 120.0, dtu=1)
    gt_dg = DeviceGaussians(ctx, random_scene(seed + 1, n, degree=1))
    gts = [render_numpy(ctx, gt_dg, c, BG) for c in cams]
    return random_scene(seed, n, degree=1, log_scale=(-4.0, 0.4)), cams, gts


def _run(ctx, sc, cams, gts, steps, graph, chunks, capacity=0, read_every=7):
    """`steps` Adam steps over a seeded view order on a fresh copy of the scene."""
    from msplat._abi import DeviceArray, TUNE_CHUNKS
    from msplat.backend import DeviceGaussians, adam_coeffs, train_step
    ctx.set_tuning(TUNE_CHUNKS, chunks)
    ctx.set_step_graph(graph, capacity)
    try:
        dg = DeviceGaussians(ctx, sc)
        targets = [DeviceArray.from_numpy(ctx, g) for g in gts]
        order = np.random.default_rng(3).integers(0, len(cams), steps)
        losses = []
        for t in range(steps):
            sizes, bc2 = adam_coeffs(LRS, t + 1)
            loss = train_step(ctx, dg, cams[order[t]], BG, targets[order[t]], sizes, bc2,
                              read_loss=(t % read_every == read_every - 1))
            if loss is not None:
                losses.append(loss)
        ctx.sync()
        f = ctx.frame()
        return dg.to_numpy(), np.array(losses, np.float32), (f.num_isects, f.num_visible, f.depth_chunks)
    finally:
        ctx.set_step_graph(False)
        ctx.set_tuning(TUNE_CHUNKS, 0)


def _same(a, b):
    for k in a:
        if k == "degree":
            continue
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


@pytest.mark.parametrize("chunks", [1, 8])
def test_replays_equal_launch_by_launch_steps(ctx, chunks):
    from msplat._abi import Context
    sc, cams, gts = _workload(ctx)
    eager = _run(ctx, sc, cams, gts, 40, False, chunks)
    gctx = Context(0)
    try:
        graph = _run(gctx, sc, cams, gts, 40, True, chunks)
        captures, replays, reruns = gctx.step_graph_stats()
    finally:
        gctx.close()
    print(f"\nK={chunks}: captures {captures}, replays {replays}, reruns {reruns}, frame {graph[2]}")
    assert captures == 1 and replays == 39 and reruns == 0
    np.testing.assert_array_equal(graph[1], eager[1])
    assert graph[2] == eager[2]  # the last frame's counts and chunk count, read after the replays
    _same(graph[0], eager[0])


def test_overflowing_replays_run_again(ctx):
    """A capacity of 2000 intersections (the frames have about 30,000): the first replay
    overflows and changes nothing, later replays in flight do the same, and the library
    runs them again launch by launch, raises the capacity and recaptures."""
    from msplat._abi import Context
    sc, cams, gts = _workload(ctx)
    eager = _run(ctx, sc, cams, gts, 24, False, 1, read_every=1000)
    gctx = Context(0)
    try:
        graph = _run(gctx, sc, cams, gts, 24, True, 1, capacity=2000, read_every=1000)
        captures, replays, reruns = gctx.step_graph_stats()
    finally:
        gctx.close()
    print(f"\ncaptures {captures}, replays {replays}, reruns {reruns}")
    assert reruns >= 1 and captures >= 2
    _same(graph[0], eager[0])


def test_trainer_with_densification_equals_launch_by_launch(ctx):
    """The trainer's schedule with densify/prune every 10 steps from step 20 (each a new
    gaussian count: a launch-by-launch step, then a new capture) and an opacity reset."""
    from msplat._abi import Context, DeviceArray, TUNE_CHUNKS
    from msplat.backend import DeviceGaussians
    from msplat.trainer import TrainConfig, Trainer
    sc, cams, gts = _workload(ctx, n=8000, seed=1200)
    out = []
    for graph in (False, True):
        c = Context(0)
        try:
            c.set_tuning(TUNE_CHUNKS, 4)
            dg = DeviceGaussians(c, sc)
            cfg = TrainConfig(iterations=200, sh_degree=1, sh_degree_interval=20, refine_every=10, warmup_length=20,
                              reset_alpha_every=4, densify_grad_thresh=2e-5, step_graph=graph)
            tr = Trainer(c, dg, cams, [DeviceArray.from_numpy(c, g) for g in gts], cfg)
            losses = [tr.step(read_loss=(i % 5 == 4)) for i in range(75)]
            c.sync()
            out.append((dg.to_numpy(), [x for x in losses if x is not None], list(tr.densify_log),
                        c.step_graph_stats()))
        finally:
            c.close()
    (e, le, de, _), (g, lg, dgl, stats) = out
    print(f"\ndensify log {de}; graph stats {stats}")
    assert de == dgl and len(de) >= 3 and any(b != a for _, b, a in de)
    assert stats[0] >= len(de) and stats[1] >= 40
    assert le == lg
    _same(g, e)
