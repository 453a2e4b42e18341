"""The training-script drop-in (aniso_amd.rasterization.rasterization, standing where
pipeline/train_gsplat.py:96-102 calls its third-party rasterizer) on the GPU.

* Its outputs are the batched differentiable path's, checked elsewhere against the CPU oracle:
  colours equal rasterize_views on the same activated parameters bit for bit, alphas are
  1 - final transmittance, info holds radii [C, N, 2] and a means2d leaf whose gradient is the
  per-view screen-space gradient.
* A training run in the script's shape (train_gsplat.py:159-190: Adam per parameter group,
  0.8 L1 + 0.2 (1 - SSIM), one random view per step) on a perturbed 200,000-Gaussian scene
  against images of the unperturbed scene: the loss falls and the PSNR to the targets rises.
Parity with the third-party rasterizer itself is not pinned (it is not importable here).
"""
import math

import numpy as np
import pytest
import torch
import torch.nn.functional as F

from aniso_amd import scenes
from aniso_amd.autograd import rasterize_views
from aniso_amd.rasterization import rasterization, viewmat_K

pytestmark = pytest.mark.gpu

W, H = 640, 360


def ssim(a, b):
    """SSIM with an 11 x 11 Gaussian window (sigma 1.5) over [1, 3, H, W] images."""
    g = torch.exp(-((torch.arange(11, device=a.device, dtype=torch.float32) - 5) ** 2) / (2 * 1.5 ** 2))
    g = g / g.sum()
    win = (g[:, None] * g[None, :]).expand(3, 1, 11, 11).contiguous()
    mu_a = F.conv2d(a, win, padding=5, groups=3)
    mu_b = F.conv2d(b, win, padding=5, groups=3)
    s_aa = F.conv2d(a * a, win, padding=5, groups=3) - mu_a ** 2
    s_bb = F.conv2d(b * b, win, padding=5, groups=3) - mu_b ** 2
    s_ab = F.conv2d(a * b, win, padding=5, groups=3) - mu_a * mu_b
    c1, c2 = 0.01 ** 2, 0.03 ** 2
    m = ((2 * mu_a * mu_b + c1) * (2 * s_ab + c2)) / ((mu_a ** 2 + mu_b ** 2 + c1) * (s_aa + s_bb + c2))
    return m.mean()


def views(count, seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        pos = rng.uniform(-0.3, 0.3, 3)
        yaw = math.radians(rng.uniform(-10, 10))
        out.append(scenes.pose_camera(pos, (math.sin(yaw), 0.0, math.cos(yaw)), (0.0, -1.0, 0.0), W, H, 60.0))
    return out


def test_drop_in_outputs(gpu):
    raw = scenes.synthetic_scene(50_000, W, H, seed=2)
    cams = views(3, seed=1)
    dev = "cuda"
    t = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float32, device=dev)
    means, quats, sh = t(raw.means), t(raw.quats), t(raw.sh)
    scales = torch.exp(t(raw.log_scales)).requires_grad_(True)
    opac = torch.sigmoid(t(raw.opacity_logits)).requires_grad_(True)
    means.requires_grad_(True)
    vms, Ks = zip(*[viewmat_K(c) for c in cams])
    img, alpha, info = rasterization(means, F.normalize(quats, dim=-1), scales, opac, sh, torch.stack(vms),
                                     torch.stack(Ks), W, H, sh_degree=3, packed=False)
    assert img.shape == (3, H, W, 3) and alpha.shape == (3, H, W, 1)
    assert info["radii"].shape == (3, raw.n, 2) and info["means2d"].shape == (3, raw.n, 2)
    assert info["n_cameras"] == 3 and info["width"] == W and info["height"] == H
    assert float(alpha.min()) >= 0.0 and float(alpha.max()) <= 1.0
    with torch.no_grad():
        ref, rinfo = rasterize_views(means, scales, F.normalize(quats, dim=-1), opac, sh, cams, info=True,
                                     activated=True)
    for v in range(3):
        assert torch.equal(img[v].detach(), ref[v])
        assert torch.equal(alpha[v, ..., 0], rinfo["alphas"][v])
    img.sum().backward()
    g = info["means2d"].grad
    assert g is not None and g.shape == (3, raw.n, 2) and torch.isfinite(g).all()
    visible = info["radii"][..., 0] > 0
    assert bool((g[~visible] == 0).all()) and bool((g[visible] != 0).any())
    assert torch.isfinite(means.grad).all() and torch.isfinite(scales.grad).all() and torch.isfinite(opac.grad).all()


def test_training_run_reduces_loss(gpu):
    """200 Adam steps in the reference script's shape on a perturbed 200k-Gaussian scene."""
    truth = scenes.synthetic_scene(200_000, W, H, seed=0)
    cams = views(8, seed=3)
    dev = "cuda"
    t = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float32, device=dev)
    mats = [viewmat_K(c) for c in cams]
    with torch.no_grad():
        targets = rasterize_views(t(truth.means), t(truth.log_scales), t(truth.quats), t(truth.opacity_logits),
                                  t(truth.sh), cams)
    rng = np.random.default_rng(1)
    n = truth.n
    params = {
        "means": torch.nn.Parameter(t(truth.means + rng.normal(0, 0.02, (n, 3)))),
        "scales": torch.nn.Parameter(t(truth.log_scales + rng.normal(0, 0.2, (n, 3)))),
        "quats": torch.nn.Parameter(t(truth.quats + rng.normal(0, 0.1, (n, 4)))),
        "opacities": torch.nn.Parameter(t(truth.opacity_logits + rng.normal(0, 0.5, n))),
        "sh0": torch.nn.Parameter(t(truth.sh[:, :1] + rng.normal(0, 0.3, (n, 1, 3)))),
        "shN": torch.nn.Parameter(t(truth.sh[:, 1:])),
    }
    lrs = {"means": 1.6e-4 * 5, "scales": 5e-3, "quats": 1e-3, "opacities": 5e-2, "sh0": 2.5e-3, "shN": 2.5e-3 / 20}
    opts = {k: torch.optim.Adam([{"params": params[k], "lr": lr, "name": k}], eps=1e-15) for k, lr in lrs.items()}

    def render(v, grad=True):
        colours = torch.cat([params["sh0"], params["shN"]], 1)
        img, alpha, info = rasterization(params["means"], F.normalize(params["quats"], dim=-1),
                                         torch.exp(params["scales"]), torch.sigmoid(params["opacities"]), colours,
                                         mats[v][0][None], mats[v][1][None], W, H, sh_degree=3, packed=False)
        return img[0]

    def mean_psnr():
        with torch.no_grad():
            return float(np.mean([-10 * math.log10(float(F.mse_loss(render(v).clamp(0, 1),
                                                                    targets[v].clamp(0, 1))))
                                  for v in range(len(cams))]))

    psnr0 = mean_psnr()
    losses = []
    for step in range(200):
        v = int(rng.integers(len(cams)))
        img = render(v)
        l1 = (img - targets[v]).abs().mean()
        loss = 0.8 * l1 + 0.2 * (1 - ssim(img.permute(2, 0, 1)[None], targets[v].permute(2, 0, 1)[None]))
        loss.backward()
        for o in opts.values():
            o.step()
            o.zero_grad(set_to_none=True)
        losses.append(float(loss))
    psnr1 = mean_psnr()
    first, last = np.mean(losses[:20]), np.mean(losses[-20:])
    print(f"training: loss {first:.4f} -> {last:.4f}, mean PSNR to targets {psnr0:.2f} -> {psnr1:.2f} dB",
          flush=True)
    assert all(math.isfinite(x) for x in losses)
    assert last < 0.8 * first
    assert psnr1 > psnr0 + 2.0
