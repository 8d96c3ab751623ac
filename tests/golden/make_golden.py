"""Writes the golden vectors in this folder from the compiled reference CPU backend
(core/cpu/msplat_cpu.cpp built by oracle/Makefile into oracle/_ref). Run where the
reference tree is mounted:

    make -C oracle && python tests/golden/make_golden.py            # every fixture
    make -C oracle && python tests/golden/make_golden.py NAME ...   # only these

Each .npz holds the inputs and the reference's outputs for one entry point of
core/metal/bindings.h. Inputs are seeded synthetic scenes.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "msplat-street_amd"))

from oracle import Reference  # noqa: E402
from msplat.camera import look_at, make_camera  # noqa: E402
from msplat.scenes import random_scene  # noqa: E402

sys.path.insert(0, os.path.dirname(HERE))
from workloads import (C1_SAMPLE, C1_STEPS, PRIOR_VARIANTS, RESIZE_CASES, UNDISTORT_CASES,  # noqa: E402
                       c1_camera_order, c1_workload, prior_workload)

BG = np.array([0.613, 0.0101, 0.3984], np.float32)
KEYS = ("means", "scales", "quats", "dc", "rest", "opac")


def cams(dtu):
    return [make_camera(look_at((2.5, 0.7, 2.9), (0.1, -0.2, 0.0)), 67, 45, 60.0, 58.0, 33.4, 22.8, dtu=dtu),
            make_camera(look_at((-3.1, -1.2, 1.4), (0.0, 0.3, -0.2)), 40, 33, 41.0, 41.0, 19.0, 17.5, dtu=dtu)]


def cam_fields(c, prefix):
    return {prefix + k: np.asarray(c[k]) for k in ("viewmat", "projmat", "cam_pos", "fx", "fy", "cx", "cy", "W", "H",
                                                   "dtu")}


def opacity_reset_vectors(ref):
    """msplat_opacity_reset (cpu:1467-1475) on logits on both sides of logit(0.2), ties
    with it, infinities and moments of both signs."""
    g = np.random.default_rng(31)
    reset = np.float32(-1.3862943611198906)
    o = np.concatenate([g.normal(0.0, 3.0, 997), [reset, np.nextafter(reset, np.float32(0)),
                                                   np.nextafter(reset, np.float32(-10)), -np.inf, np.inf, 0.0,
                                                   -0.0]]).astype(np.float32)
    m1 = g.normal(0, 1, o.size).astype(np.float32)
    m2 = np.abs(g.normal(0, 1, o.size)).astype(np.float32)
    d = dict(o=o.copy(), m1=m1.copy(), m2=m2.copy(), reset=reset)
    ref.opacity_reset(o, m1, m2, reset)
    d.update(out_o=o, out_m1=m1, out_m2=m2)
    return d


def trajectory_c1_vectors(ref):
    """300 Adam steps of Model::fullIteration (model.cpp:626-682) over the reference's
    msplat_train_step (cpu:1204-1327), with Model's schedule: Adam scalars per step
    (model.cpp:638-651), the means LR decayed after each step over a 300-step schedule
    (model.cpp:140-143), SH degree 1 in use from step 1 (sh_degree_interval 1,
    model.cpp:535), no densification. Keeps the per-step loss, the final parameters of a
    seeded sample of Gaussians, per-group sums, and the final PSNR of every view."""
    import hashlib

    from oracle import Oracle
    orc = Oracle()  # for the schedule's scalars only (pinned against model.cpp elsewhere)
    sc, gt_scene, cs = c1_workload()
    gts = [ref.render(gt_scene, c, BG)[0] for c in cs]
    state = {k: sc[k].copy() for k in KEYS}
    for k in KEYS:
        state["m1_" + k] = np.zeros_like(sc[k])
        state["m2_" + k] = np.zeros_like(sc[k])
    N = sc["means"].shape[0]
    state.update(degree=1, vis=np.zeros(N, np.float32), gn=np.zeros(N, np.float32), ms=np.zeros(N, np.float32))
    lrs = [1.6e-4, 5e-3, 1e-3, 2.5e-3, 1.25e-4, 5e-2]
    order = c1_camera_order(C1_STEPS)
    losses = []
    for t in range(1, C1_STEPS + 1):
        sc_bc = [orc.adam_scalars(np.float32(lr), t) for lr in lrs]
        steps = np.array([s for s, _ in sc_bc], np.float32)
        bc2s = np.array([b for _, b in sc_bc], np.float32)
        loss, _ = ref.train_step(state, cs[order[t - 1]], BG, gts[order[t - 1]], steps, bc2s)
        losses.append(loss)
        lrs[0] = float(orc.means_lr(t, C1_STEPS))
    final = dict(state, degree=1)
    psnr = np.array([ref.image_metrics(ref.render(final, c, BG)[0], g)[0] for c, g in zip(cs, gts)])
    sample = np.sort(np.random.default_rng(5).choice(N, C1_SAMPLE, replace=False)).astype(np.int32)
    d = dict(order=order, losses=np.array(losses, np.float32), psnr=psnr, sample=sample, bg=BG,
             gt_sha256=np.frombuffer(hashlib.sha256(np.ascontiguousarray(np.stack(gts)).tobytes()).digest(),
                                     np.uint8))
    for k in KEYS:
        d["final_" + k] = state[k][sample]
        d["sum_" + k] = np.float64(state[k].astype(np.float64).sum())
        d["sumsq_" + k] = np.float64((state[k].astype(np.float64) ** 2).sum())
    return d


def image_pipeline_vectors(ref):
    """msplat_resize_area (cpu:1609-1639) on RGBA8 and float sources, down, up and at
    equal size; msplat_undistort (cpu:1641-1669) with barrel and pincushion distortion,
    crops inside and outside the source, and no distortion."""
    g = np.random.default_rng(41)
    d = {}
    for i, (kind, sw, sh, dw, dh) in enumerate(RESIZE_CASES):
        src = (g.integers(0, 256, (sh, sw, 4), dtype=np.uint8) if kind == "u8"
               else g.random((sh, sw, 3), dtype=np.float32))
        d[f"resize_src{i}"] = src
        d[f"resize_out{i}"] = ref.resize_area(src, dw, dh)
    src = g.random((60, 80, 3), dtype=np.float32)
    d["undistort_src"] = src
    for i, (intr, dist, rx, ry, dw, dh) in enumerate(UNDISTORT_CASES):
        d[f"undistort_out{i}"] = ref.undistort(src, np.array(intr, np.float32), np.array(dist, np.float32), rx, ry,
                                               dw, dh)
    return d


def scene_file_vectors(ref):
    """saveGaussianPly, saveGaussianSplat and loadGaussianPly (save_gaussians.cpp:10-191)
    of the reference on a 120-Gaussian SH degree 3 scene with non-unit quaternions and
    two equal splat sort weights, plain and with the dataset transform (scale 2.5,
    translation (10, -3, 1.5)), and a degree 0 scene: the written files' bytes, and what
    the reader makes of the transformed PLY."""
    import tempfile
    sc = random_scene(51, 120, degree=3, log_scale=(-3.0, 0.6))
    sc["quats"] = (sc["quats"] * np.random.default_rng(52).uniform(0.5, 1.5, (120, 1))).astype(np.float32)
    sc["scales"][7], sc["opac"][7] = sc["scales"][3], sc["opac"][3]  # a tie in the .splat order
    sc0 = random_scene(53, 40, degree=0, log_scale=(-3.0, 0.6))
    sc0["rest"] = np.zeros((40, 0, 3), np.float32)
    d = {k: sc[k] for k in KEYS}
    d.update({"deg0_" + k: sc0[k] for k in KEYS})
    tr = (10.0, -3.0, 1.5)
    with tempfile.TemporaryDirectory() as tmp:
        def written(name, scene, **kw):
            path = os.path.join(tmp, name)
            ref.save(scene, path, **kw)
            d[name.replace(".", "_")] = np.frombuffer(open(path, "rb").read(), np.uint8)
            return path
        written("plain.ply", sc, step=7000)
        crs = written("crs.ply", sc, step=30000, scale=2.5, translation=tr, keep_crs=True)
        written("plain.splat", sc, splat=True)
        written("crs.splat", sc, scale=2.5, translation=tr, keep_crs=True, splat=True)
        written("deg0.ply", sc0, step=5)
        back = ref.load_ply(crs, 120, 15, scale=2.5, translation=tr, keep_crs=True)
    d.update({"loaded_" + k: back[k] for k in KEYS})
    d["loaded_step"] = np.int32(back["step"])
    return d


def prior_step_vectors(ref):
    """msplat_train_step with a PriorStep (cpu:1204-1327): the prior-guided case of
    tests/workloads.py with every feature (aux rasterizer, learned sky, depth prior, sky
    mask, fill, exposure, needle cap, unkept pixels), one step from zero moments. The
    inputs are regenerated from their seed; the outputs are stored."""
    sc, cam, gt, prior = prior_workload(seed=3, features=PRIOR_VARIANTS["all"])
    st = {k: np.ascontiguousarray(sc[k], np.float32).copy() for k in KEYS}
    for k in KEYS:
        st["m1_" + k] = np.zeros_like(st[k])
        st["m2_" + k] = np.zeros_like(st[k])
    n = st["means"].shape[0]
    st.update(degree=sc["degree"], vis=np.zeros(n, np.float32), gn=np.zeros(n, np.float32), ms=np.zeros(n, np.float32))
    steps = np.array([1.6e-4, 5e-3, 1e-3, 2.5e-3, 1.25e-4, 5e-2], np.float32)
    bc2 = np.full(6, np.float32(np.sqrt(np.float32(1 - 0.999))), np.float32)
    loss, radii, depth, fT = ref.train_step_prior(st, cam, BG, gt, steps, bc2, prior)
    d = {"loss": np.float32(loss), "depth": depth, "final_T": fT, "radii": radii}
    d.update({"out_" + k: st[k] for k in KEYS})
    d.update({"out_" + k: prior[k] for k in ("sky_tex", "expo_params", "loss_terms")})
    return d


EXTRA = {"opacity_reset": opacity_reset_vectors, "trajectory_c1": trajectory_c1_vectors,
         "image_pipeline": image_pipeline_vectors, "scene_files": scene_file_vectors,
         "prior_step": prior_step_vectors}


def main(only=()):
    ref = Reference()
    if only:
        for name in only:
            d = EXTRA[name](ref)
            np.savez_compressed(os.path.join(HERE, name + ".npz"), **d)
            print(name, os.path.getsize(os.path.join(HERE, name + ".npz")))
        return
    out = {name: fn(ref) for name, fn in EXTRA.items()}
    # render at SH degrees 0, 1, 3 from two cameras (non-tile-multiple sizes)
    for deg in (0, 1, 3):
        sc = random_scene(100 + deg, 400, degree=deg, log_scale=(-3.0, 0.6))
        d = {k: sc[k] for k in KEYS}
        d["degree"] = np.int32(deg)
        for i, c in enumerate(cams(deg)):
            img, fT = ref.render(sc, c, BG)
            d.update(cam_fields(c, f"cam{i}_"))
            d[f"img{i}"], d[f"final_T{i}"] = img, fT
        d["bg"] = BG
        out[f"render_deg{deg}"] = d

    # three training steps (SH degree 3 in use), zero-initialised moments
    sc = random_scene(7, 400, degree=3, log_scale=(-3.0, 0.6))
    state = {k: sc[k].copy() for k in KEYS}
    for k in KEYS:
        state["m1_" + k] = np.zeros_like(sc[k])
        state["m2_" + k] = np.zeros_like(sc[k])
    N = sc["means"].shape[0]
    state.update(degree=3, vis=np.zeros(N, np.float32), gn=np.zeros(N, np.float32), ms=np.zeros(N, np.float32))
    d = {"init_" + k: sc[k].copy() for k in KEYS}
    lrs = np.array([1.6e-4, 5e-3, 1e-3, 2.5e-3, 1.25e-4, 5e-2], np.float32)
    cs = cams(3)
    rng = np.random.default_rng(11)
    losses, steps_all, bc2_all = [], [], []
    for t in range(3):
        c = cs[t % 2]
        gt = np.clip(rng.random((c["H"], c["W"], 3)), 0, 1).astype(np.float32)
        bc1 = np.float32(1) - np.float32(0.9) ** np.float32(t + 1)
        bc2 = np.float32(1) - np.float32(0.999) ** np.float32(t + 1)
        steps = (lrs / bc1).astype(np.float32)
        bc2s = np.full(6, np.sqrt(bc2), np.float32)
        loss, radii = ref.train_step(state, c, BG, gt, steps, bc2s)
        losses.append(loss)
        steps_all.append(steps)
        bc2_all.append(bc2s)
        d[f"gt{t}"] = gt
        d[f"radii{t}"] = radii
    for i, c in enumerate(cs):
        d.update(cam_fields(c, f"cam{i}_"))
    d.update(bg=BG, losses=np.array(losses, np.float32), steps=np.array(steps_all), bc2s=np.array(bc2_all))
    for k in KEYS:
        for pre in ("", "m1_", "m2_"):
            d["out_" + pre + k] = state[pre + k]
    for k in ("vis", "gn", "ms"):
        d["out_" + k] = state[k]
    out["train_steps"] = d

    # densify: one call with splits, duplicates and culls, huge-culling on and off
    g = np.random.default_rng(21)
    N, fr = 300, 9
    cap = 3 * N
    widths = [3, 3, 4, 3, fr, 1]
    inputs = []
    for i in range(18):
        a = g.normal(0, 1, (N, widths[i % 6])).astype(np.float32)
        if i == 1:
            a = g.normal(-3.5, 1.5, (N, 3)).astype(np.float32)
        if i >= 6:
            a = (np.abs(a) * 0.01).astype(np.float32)
        inputs.append(a)
    vc = g.integers(0, 4, N).astype(np.float32)
    gn = (g.random(N) * 6e-3 * vc).astype(np.float32)
    ms = (g.random(N) * 0.3).astype(np.float32)
    d = {f"in{i}": a for i, a in enumerate(inputs)}
    d.update(vc=vc, gn=gn, ms=ms, N=np.int32(N), fr=np.int32(fr), seed=np.uint32((700 * 2654435761) % 2 ** 32))
    for ch in (0, 1):
        bufs = []
        for i in range(18):
            b = np.zeros((cap, widths[i % 6]), np.float32)
            b[:N] = inputs[i]
            bufs.append(b)
        n = ref.densify(bufs, N, cap, gn, vc, ms, 40.0, fr, int(d["seed"]), cull_alpha=0.1, check_huge=ch)
        d[f"count_huge{ch}"] = np.int32(n)
        for i in range(18):
            d[f"out{ch}_{i}"] = bufs[i][:n]
    out["densify"] = d

    # kNN and Gaussian init on a street-like cloud with duplicates
    road = np.c_[g.uniform(-20, 20, 600), g.normal(0, 0.02, 600), g.uniform(0, 80, 600)]
    facade = np.c_[8 + g.normal(0, 0.05, 300), g.uniform(0, 10, 300), g.uniform(0, 80, 300)]
    pts = np.vstack([road, facade, g.normal(0, 300, (60, 3)), np.repeat(g.uniform(-1, 1, (5, 3)), 4, 0)]).astype(
        np.float32)
    rgb = g.integers(0, 256, (pts.shape[0], 3), dtype=np.uint8)
    s, q, dc, o = ref.init_gaussians(pts, rgb, seed=42)
    out["knn_init"] = dict(points=pts, rgb=rgb, knn=ref.knn3(pts), scales=s, quats=q, dc=dc, opac=o)

    # radix sort (stable, low key_bits), metrics, RGBA8 packing
    d = {}
    for i, (n, bits, hi) in enumerate([(1, 32, 10), (1025, 32, 7), (3000, 32, 2 ** 30), (2049, 64, 3),
                                       (1500, 45, 2 ** 45)]):
        keys = g.integers(0, hi, n, dtype=np.uint64)
        vals = np.arange(n, dtype=np.int32)
        k2, v2 = ref.radix_sort(keys, vals, bits)
        d.update({f"keys{i}": keys, f"bits{i}": np.int32(bits), f"sorted{i}": k2, f"order{i}": v2})
    out["radix_sort"] = d
    a = g.random((37, 53, 3), dtype=np.float32)
    b = np.clip(a + g.normal(0, 0.05, a.shape), 0, 1).astype(np.float32)
    img = g.uniform(-0.2, 1.2, (21, 17, 3)).astype(np.float32)
    out["image_ops"] = dict(a=a, b=b, metrics=ref.image_metrics(a, b), img=img, rgba=ref.pack_rgba8(img))

    for name, d in out.items():
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **d)
        print(name, os.path.getsize(os.path.join(HERE, name + ".npz")))


if __name__ == "__main__":
    main(tuple(sys.argv[1:]))
