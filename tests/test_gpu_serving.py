"""Serving with continuous batching on the GPU: every request's RGBA8 image is the same
alone (one msplat_hip_render per request) and in the mix (batched with other requests
of its scene, two slots in flight at once), and admission keeps each batch within the
slot's share of the HBM budget."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BG = (0.1, 0.2, 0.3)
SMALL = [((64, 48), 0.4), ((128, 72), 0.3), ((96, 96), 0.2), ((50, 70), 0.1)]


def _scenes(ctx):
    from msplat.backend import DeviceGaussians
    from msplat.scenes import random_scene
    out, centers = [], [(0.0, 0.0, 0.0), (20.0, 0.0, 5.0)]
    for k, (n, c) in enumerate(zip((3000, 5000), centers)):
        sc = random_scene(40 + k, n, degree=3, mean_range=1.2, log_scale=(-3.0, 0.4))
        sc["means"] = (sc["means"] + np.asarray(c, np.float32)).astype(np.float32)
        out.append(DeviceGaussians(ctx, sc))
    ctx.sync()
    return out, centers


def test_results_are_the_same_alone_and_in_the_mix(ctx):
    from msplat.serving import Server, poisson_requests, serve_plain
    scenes, centers = _scenes(ctx)
    reqs = poisson_requests(4000.0, 0.012, centers, seed=3, radius=(3.0, 6.0), resolutions=SMALL)
    assert len(reqs) >= 30
    srv = Server(scenes, slots=2, max_views=8, background=BG, keep_images=True)
    try:
        done = srv.run(reqs)
    finally:
        srv.close()
    assert len(done) == len(reqs) and not srv.rejected
    assert max(r.batch_size for r in done) > 1  # requests were batched
    alone = serve_plain(scenes, [type(r)(r.rid, r.scene, r.cam, 0.0) for r in reqs], background=BG,
                        keep_images=True, depth_chunks=1)
    by_id = {r.rid: r for r in alone}
    for r in done:
        np.testing.assert_array_equal(r.image, by_id[r.rid].image, err_msg=f"request {r.rid}")
    assert any(r.image[..., :3].std() > 2 for r in done)  # scenes are visible
    # the plain form's default depth-chunked frames: the same up to float rounding, at
    # most one RGBA8 step on a few pixels
    chunked = serve_plain(scenes, [type(r)(r.rid, r.scene, r.cam, 0.0) for r in reqs], background=BG,
                          keep_images=True)
    diffs = [np.abs(r.image.astype(int) - by_id[r.rid].image.astype(int)) for r in chunked]
    assert max(int(d.max()) for d in diffs) <= 1
    assert np.mean([np.mean(d > 0) for d in diffs]) < 1e-3


def test_admission_keeps_batches_within_the_budget(ctx):
    from msplat.serving import Request, Server, poisson_requests
    scenes, centers = _scenes(ctx)
    reqs = poisson_requests(4000.0, 0.006, centers, seed=5, radius=(3.0, 6.0), resolutions=[((96, 96), 1.0)])
    resident = sum(s.nbytes for s in scenes)
    probe = Server(scenes, slots=1, background=BG)
    one = max(probe.working_set([r]) for r in reqs)
    probe.close()
    # each slot's share admits one 96x96 view but not two, and never a 1920x1080 one
    srv = Server(scenes, slots=2, max_views=8, hbm_budget=resident + 2 * 1.5 * one, background=BG, adapt=False)
    from msplat.camera import look_at, make_camera
    big = Request(len(reqs), 0, make_camera(look_at((0, 0, 5.0), (0, 0, 0)), 1920, 1080, 1500.0, 1500.0, dtu=3),
                  0.001)
    try:
        done = srv.run(reqs + [big])
    finally:
        srv.close()
    assert [r.rid for r in srv.rejected] == [big.rid]
    assert len(done) == len(reqs) and all(r.batch_size == 1 for r in done)


def test_scenes_admitted_and_evicted_while_serving(ctx, tmp_path):
    """A scene written to a PLY file and loaded back is admitted into a running server
    (add_scene), served beside the resident one with the images it gives alone, evicted
    from another thread part-way through a burst (remove_scene: its later requests are
    refused, the other scene's served), and a scene too large for the budget is refused."""
    import threading
    import time
    from msplat.scene_io import load_ply, save_ply
    from msplat.serving import MIN_VIEW, Server, poisson_requests, serve_plain
    scenes, centers = _scenes(ctx)
    path = tmp_path / "scene1.ply"
    save_ply(ctx, scenes[1], path)
    loaded, _ = load_ply(ctx, path, capacity=scenes[1].n)
    srv = Server(scenes[:1], slots=2, max_views=8, background=BG, keep_images=True)
    try:
        idx = srv.add_scene(loaded)
        assert idx == 1 and srv.resident == scenes[0].nbytes + loaded.nbytes
        reqs = poisson_requests(2000.0, 0.3, centers, seed=8, radius=(3.0, 6.0), resolutions=SMALL)
        evicted = []

        def evict():
            srv.remove_scene(idx)
            evicted.append(srv._now())

        timer = threading.Timer(0.15, evict)
        timer.start()
        done = srv.run(reqs)
        timer.join()
        t_evict = evicted[0]
        assert srv.scenes[idx] is None and srv.resident == scenes[0].nbytes
        mine = [r for r in done if r.scene == idx]
        refused = [r for r in srv.rejected if r.scene == idx]
        assert mine and refused and not [r for r in srv.rejected if r.scene == 0]
        assert len(mine) + len(refused) == sum(r.scene == idx for r in reqs)
        assert all(r.arrival >= t_evict - 0.05 for r in refused)  # refused only around and after the eviction
        alone = serve_plain([scenes[0], scenes[1]], [type(r)(r.rid, r.scene, r.cam, 0.0) for r in done],
                            background=BG, keep_images=True, depth_chunks=1)
        by_id = {r.rid: r for r in alone}
        for r in done:
            np.testing.assert_array_equal(r.image, by_id[r.rid].image, err_msg=f"request {r.rid}")
        with pytest.raises(ValueError):
            srv.remove_scene(idx)
    finally:
        srv.close()
    # a budget one byte short of the resident scenes plus each slot's smallest view
    small = Server(scenes[:1], slots=1, background=BG)
    need = small._working_set(0, 1, MIN_VIEW[0] * MIN_VIEW[1])
    small.close()
    tight = Server(scenes[:1], slots=2, background=BG, hbm_budget=scenes[0].nbytes + 2 * need + scenes[1].nbytes - 1)
    try:
        assert tight.add_scene(scenes[1]) is None and tight.resident == scenes[0].nbytes
    finally:
        tight.close()
    roomy = Server(scenes[:1], slots=2, background=BG, hbm_budget=scenes[0].nbytes + 4 * need + scenes[1].nbytes)
    try:
        assert roomy.add_scene(scenes[1]) is not None
    finally:
        roomy.close()


def test_requests_queued_past_the_deadline_are_shed(ctx):
    """Load shedding: two hundred requests that all arrive at once, with a 2 ms queueing
    bound. The two slots cannot start them all within 2 ms, so the rest are shed; every
    served request started rendering within the bound (plus the host's scheduling slack),
    and every request is either served or shed. Without the bound all are served."""
    from msplat.serving import Request, Server, poisson_requests
    scenes, centers = _scenes(ctx)
    base = poisson_requests(4000.0, 0.05, centers, seed=5, radius=(3.0, 6.0), resolutions=SMALL)
    reqs = [Request(r.rid, r.scene, r.cam, 0.0) for r in base]
    assert len(reqs) >= 150
    srv = Server(scenes, slots=2, max_views=8, background=BG, max_wait_ms=2.0)
    try:
        done = srv.run(reqs)
    finally:
        srv.close()
    assert len(done) + len(srv.shed) == len(reqs) and not srv.rejected
    assert done and srv.shed
    assert max(r.launched - r.arrival for r in done) <= 0.002 + 0.05
    everything = Server(scenes, slots=2, max_views=8, background=BG)
    try:
        assert len(everything.run([Request(r.rid, r.scene, r.cam, 0.0) for r in base])) == len(reqs)
        assert not everything.shed
    finally:
        everything.close()
