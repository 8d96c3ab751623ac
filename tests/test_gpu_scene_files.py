"""Trained-scene files written and read by the HIP backend: PLY and .splat exports byte for
byte the reference's writers' (tests/golden/scene_files.npz, from save_gaussians.cpp), the
reference's PLY read back to the reference reader's parameters, and the .msplat v2
checkpoint checked field by field by an independent reader of its layout
(model.cpp:282-370) and by resuming training from it."""
import os
import struct

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu

KEYS = ("means", "scales", "quats", "dc", "rest", "opac")
TR = (10.0, -3.0, 1.5)


def _scene(d, prefix="", degree=3):
    return dict({k: d[prefix + k] for k in KEYS}, degree=degree)


def test_exports_equal_reference_bytes(ctx, tmp_path):
    from msplat.backend import DeviceGaussians
    from msplat.scene_io import save_ply, save_splat
    d = golden("scene_files")
    dg = DeviceGaussians(ctx, _scene(d), capacity=200)
    cases = [("plain.ply", lambda p: save_ply(ctx, dg, p, step=7000)),
             ("crs.ply", lambda p: save_ply(ctx, dg, p, step=30000, scale=2.5, translation=TR, keep_crs=True)),
             ("plain.splat", lambda p: save_splat(ctx, dg, p)),
             ("crs.splat", lambda p: save_splat(ctx, dg, p, scale=2.5, translation=TR, keep_crs=True))]
    for name, write in cases:
        path = str(tmp_path / name)
        write(path)
        got = np.frombuffer(open(path, "rb").read(), np.uint8)
        np.testing.assert_array_equal(got, d[name.replace(".", "_")], err_msg=name)
    d0 = DeviceGaussians(ctx, _scene(d, "deg0_", 0))
    path = str(tmp_path / "deg0.ply")
    save_ply(ctx, d0, path, step=5)
    np.testing.assert_array_equal(np.frombuffer(open(path, "rb").read(), np.uint8), d["deg0_ply"])


def test_reference_ply_loads(ctx, tmp_path):
    """The reference's transformed PLY read back with the transform: the reference
    reader's parameters bit for bit; its 'comment msplat v30000' is not the comment the
    reader parses, so the step is 0 there and here."""
    from msplat.scene_io import load_ply, ply_info
    d = golden("scene_files")
    path = str(tmp_path / "crs.ply")
    open(path, "wb").write(d["crs_ply"].tobytes())
    assert ply_info(path) == {"num_points": 120, "rest_bases": 15, "degree": 3, "step": 0}
    dg, step = load_ply(ctx, path, scale=2.5, translation=TR, keep_crs=True)
    assert step == int(d["loaded_step"]) == 0 and dg.n == 120 and dg.capacity == 480
    got = dg.to_numpy()
    for k in KEYS:
        np.testing.assert_array_equal(got[k].reshape(d["loaded_" + k].shape), d["loaded_" + k], err_msg=k)
    open(str(tmp_path / "bad.ply"), "wb").write(b"not a ply\n")
    from msplat._abi import MsplatError
    with pytest.raises(MsplatError):
        load_ply(ctx, str(tmp_path / "bad.ply"))


def _read_checkpoint(path):
    """The .msplat layout, read independently: magic, version, step, count, SH degree,
    Adam step count (u32 each), six learning rates, the means' initial and final rates
    (f32), eighteen tensors (u32 ndim, i64 shape, u64 bytes, f32 data), then the sky and
    exposure flags (u32)."""
    b = open(path, "rb").read()
    magic, version, step, n, deg, adam_t = struct.unpack_from("<6I", b, 0)
    lrs = struct.unpack_from("<8f", b, 24)
    off, tensors = 56, []
    for _ in range(18):
        ndim, = struct.unpack_from("<I", b, off)
        shape = struct.unpack_from(f"<{ndim}q", b, off + 4)
        nbytes, = struct.unpack_from("<Q", b, off + 4 + 8 * ndim)
        off += 12 + 8 * ndim
        tensors.append(np.frombuffer(b, np.float32, nbytes // 4, off).reshape(shape))
        off += nbytes
    sky, expo = struct.unpack_from("<2I", b, off)
    assert off + 8 == len(b)
    return dict(magic=magic, version=version, step=step, n=n, degree=deg, adam_t=adam_t, lrs=lrs, tensors=tensors,
                sky=sky, expo=expo)


def test_checkpoint_layout_and_resume(ctx, tmp_path):
    """Three train steps, a checkpoint, two more steps; a trainer resumed from the
    checkpoint takes the same two steps to the same parameters and moments, bit for bit."""
    from msplat._abi import DeviceArray
    from msplat.backend import DeviceGaussians
    from msplat.scene_io import checkpoint_info
    from msplat.scenes import random_scene, ring_cameras
    from msplat.trainer import TrainConfig, Trainer
    sc = random_scene(61, 3000, degree=3, log_scale=(-3.5, 0.5))
    cams = ring_cameras(4, 4.0, 96, 80, 90.0, 90.0)
    rng = np.random.default_rng(62)
    tg = [DeviceArray.from_numpy(ctx, rng.uniform(0, 1, (80, 96, 3)).astype(np.float32)) for _ in cams]
    cfg = TrainConfig(iterations=100, sh_degree=3, sh_degree_interval=1, densify=False)
    a = Trainer(ctx, DeviceGaussians(ctx, sc, capacity=5000), cams, tg, cfg)
    for _ in range(3):
        a.step()
    path = str(tmp_path / "run.msplat")
    a.save_checkpoint(path)
    f = _read_checkpoint(path)
    assert (f["magic"], f["version"], f["step"], f["n"], f["degree"], f["adam_t"]) == (0x4C50534D, 2, 3, 3000, 3, 3)
    assert f["lrs"][:6] == tuple(np.float32(v) for v in a.lr) and f["sky"] == 0 and f["expo"] == 0
    st = a.dg.to_numpy()
    for i, k in enumerate(KEYS):
        for j, pre in enumerate(("", "m1_", "m2_")):
            t = f["tensors"][6 * j + i]
            want = st[pre + k].reshape(3000, 15, 3) if k == "rest" else st[pre + k].reshape(3000, -1)
            np.testing.assert_array_equal(t, want, err_msg=pre + k)
    assert checkpoint_info(path)["num_points"] == 3000
    b = Trainer.from_checkpoint(ctx, path, cams, tg, cfg)
    b.order, b.order_pos = list(a.order), a.order_pos  # the same view sequence from here on
    b.rng.bit_generator.state = a.rng.bit_generator.state
    assert b.dg.capacity == 12000 and b.step_count == 3 and b.adam_step_count == 3
    for _ in range(2):
        a.step()
        b.step()
    sa, sb = a.dg.to_numpy(), b.dg.to_numpy()
    for k in KEYS:
        for pre in ("", "m1_", "m2_"):
            np.testing.assert_array_equal(sb[pre + k], sa[pre + k], err_msg=pre + k)
    assert os.path.getsize(path) > 3000 * 59 * 4 * 3
