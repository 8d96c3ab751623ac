"""Host-side logic of the serving loop: the Poisson request generator and the latency
statistics (no GPU)."""
import numpy as np

from msplat.serving import RESOLUTIONS, Request, latency_stats, poisson_requests


def test_poisson_requests_follow_the_mix():
    centers = [(0.0, 0.0, 0.0), (50.0, 0.0, 0.0), (0.0, 50.0, 0.0), (0.0, 0.0, 50.0)]
    reqs = poisson_requests(1000.0, 4.0, centers, seed=1)
    assert 3800 < len(reqs) < 4200  # 4000 expected, sd 63
    gaps = np.diff([r.arrival for r in reqs])
    assert gaps.min() > 0 and abs(gaps.mean() - 1e-3) < 1e-4
    counts = {}
    for r in reqs:
        counts[(r.cam["W"], r.cam["H"])] = counts.get((r.cam["W"], r.cam["H"]), 0) + 1
    for size, w in RESOLUTIONS:
        assert abs(counts[size] / len(reqs) - w) < 0.03
    for r in reqs:
        c = np.asarray(centers[r.scene])
        eye = r.cam["cam_to_world"][:3, 3]
        assert 3.0 - 1e-4 <= np.linalg.norm(eye - c) <= 10.0 + 1e-4
        # the camera looks at the scene center: it projects to the principal point
        p = r.cam["viewmat"].reshape(4, 4) @ np.r_[c, 1.0]
        assert p[2] > 0 and abs(p[0] / p[2]) < 1e-4 and abs(p[1] / p[2]) < 1e-4
    assert {r.scene for r in reqs} == {0, 1, 2, 3}


def test_latency_stats():
    cam = {"W": 1920, "H": 1080}
    reqs = []
    for i in range(100):
        r = Request(i, 0, cam, arrival=0.01 * i)
        r.done = r.arrival + 0.001 * (i + 1)  # latencies 1..100 ms
        r.batch_size = 2
        reqs.append(r)
    st = latency_stats(reqs)
    assert st["completed"] == 100 and st["mean_batch"] == 2
    assert abs(st["p50_ms"] - 50.5) < 1e-6 and abs(st["p99_ms"] - 99.01) < 1e-6
    assert abs(st["views_per_s"] - 100 / (0.99 + 0.1)) < 1e-2
    assert st["by_resolution"]["1920x1080"]["n"] == 100
