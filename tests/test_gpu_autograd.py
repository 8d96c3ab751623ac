"""torch.autograd.Function over the HIP kernels (msplat.autograd.rasterize), checked against
central finite differences (h = 1e-3) of the FP64 CPU renderer: 64 Gaussians, 32x32 image,
every parameter of every Gaussian, relative tolerance 1e-3. The FP32 gradients come from
the HIP raster backward and projection backward; the loss is differentiated by torch in
float64."""
import numpy as np
import pytest

import gradcheck_scene as gs

pytestmark = pytest.mark.gpu


def _torch_train_loss(img, gt, w=gs.SSIM_WEIGHT):
    import torch
    import torch.nn.functional as F
    g1 = torch.tensor(gs.GAUSS, dtype=torch.float64)
    k = (g1[:, None] * g1[None, :])[None, None].repeat(3, 1, 1, 1)
    x = torch.tensor(gt, dtype=torch.float64).permute(2, 0, 1)[None]
    y = img.double().permute(2, 0, 1)[None]
    blur = lambda a: F.conv2d(a, k, padding=5, groups=3)
    mx, my = blur(x), blur(y)
    vx, vy, cxy = blur(x * x) - mx * mx, blur(y * y) - my * my, blur(x * y) - mx * my
    C1, C2 = 1e-4, 9e-4
    ssim = ((2 * mx * my + C1) * (2 * cxy + C2)) / ((mx * mx + my * my + C1) * (vx + vy + C2))
    return w * (1 - ssim.mean()) + (1 - w) * (x - y).abs().mean()


@pytest.mark.parametrize("view_dependent,loss", [(False, "linear"), (False, "train"), (True, "linear"),
                                                 (True, "train")])
def test_autograd_matches_fp64_finite_differences(oracle, view_dependent, loss):
    import torch
    from msplat.autograd import rasterize
    sc, cam = gs.scene(view_dependent), gs.camera()
    dev = torch.device("cuda:0")
    params = [torch.tensor(sc[k], device=dev, requires_grad=True) for k in gs.KEYS]
    img = rasterize(*params, cam, tuple(gs.BG))
    # the FP32 image against the FP64 renderer
    ref = oracle.render(sc, cam, gs.BG, dtype=np.float64)[2]["img"]
    assert np.abs(img.detach().cpu().numpy() - ref).max() < 2e-6
    if loss == "linear":
        L = (img.double().cpu() * torch.tensor(gs.linear_weights())).sum()
    else:
        L = _torch_train_loss(img.cpu(), gs.target(oracle, sc, cam))
    L.backward()
    grads = {k: p.grad.cpu().numpy() for k, p in zip(gs.KEYS, params)}
    fds = gs.finite_differences(oracle, sc, cam, loss)
    checked = gs.compare(grads, fds, f"{loss} view_dependent={view_dependent}")
    assert checked >= 64 * 20


def test_autograd_backward_twice_and_interleaved():
    """backward re-renders its own frame: two graphs built before either backward give
    the gradients each would give alone."""
    import torch
    from msplat.autograd import rasterize
    dev = torch.device("cuda:0")
    cam = gs.camera()
    R = torch.tensor(gs.linear_weights(), dtype=torch.float32, device=dev)
    outs = []
    for vd in (False, True):
        sc = gs.scene(vd, seed=11)
        ps = [torch.tensor(sc[k], device=dev, requires_grad=True) for k in gs.KEYS]
        outs.append((ps, (rasterize(*ps, cam, tuple(gs.BG)) * R).sum()))
    for ps, L in reversed(outs):
        L.backward()
    for vd, (ps, _) in zip((False, True), outs):
        sc = gs.scene(vd, seed=11)
        qs = [torch.tensor(sc[k], device=dev, requires_grad=True) for k in gs.KEYS]
        (rasterize(*qs, cam, tuple(gs.BG)) * R).sum().backward()
        for p, q in zip(ps, qs):
            torch.testing.assert_close(p.grad, q.grad, rtol=0, atol=0)
