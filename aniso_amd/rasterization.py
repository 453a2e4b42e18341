"""A drop-in for the rasterizer call of the reference's training script
(pipeline/train_gsplat.py:96-102):

    img, alpha, info = rasterization(means, F.normalize(quats, dim=-1), torch.exp(scales),
                                     torch.sigmoid(opacities), colours, viewmat[None], K[None],
                                     w, h, sh_degree=sh_degree, packed=False)

with the arguments that call passes: activated parameters (unit quaternions are normalized
again inside, scales are standard deviations, opacities in [0, 1]), SH colour coefficients
[N, K, 3], C world-to-camera matrices [C, 4, 4] and intrinsics [C, 3, 3] (pinhole, +x right,
+y down, +z forward), one image size. All C cameras are rendered by one batched pass of the
hand-written HIP kernels (ans_diff_views_*), forward and backward.

Returns (colours [C, H, W, 3], alphas [C, H, W, 1], info). info carries what the script's
densification strategy reads (train_gsplat.py:177, :185): "means2d" ([C, N, 2] leaf whose .grad
is dL/d(pixel centre) after backward), "radii" ([C, N, 2] int32 pixels, 0 where culled; this
renderer's splat radius is isotropic, so both entries are equal), "width", "height",
"n_cameras", "packed" (False).

Conventions are the reference renderer's (src/render.cpp:39-104, :193-221): near plane 0.2,
3-sigma radius, 0.3 low-pass, alpha clamped to 0.99, skip below 1/255, stop below T = 1e-4, SH
colour + 0.5 clamped at zero, black background. The alphas carry no gradient. Parity with the
third-party rasterizer itself is not pinned (it is not importable here); the colours and
gradients are checked against this project's CPU oracle.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from .autograd import rasterize_views
from .renderer import RenderOptions
from .scenes import Camera

SH_C0 = 0.28209479177387814


def _cameras(viewmats, Ks, width, height):
    vm = viewmats.detach().double().cpu().numpy()
    ks = Ks.detach().double().cpu().numpy()
    if vm.ndim != 3 or vm.shape[1:] != (4, 4) or ks.shape != (vm.shape[0], 3, 3):
        raise ValueError("viewmats must be [C, 4, 4] and Ks [C, 3, 3]")
    return [Camera(vm[c, :3, :3], vm[c, :3, 3], ks[c, 0, 0], ks[c, 1, 1], ks[c, 0, 2], ks[c, 1, 2], int(width),
                   int(height)) for c in range(vm.shape[0])]


def rasterization(means, quats, scales, opacities, colors, viewmats, Ks, width, height, sh_degree=None,
                  packed=False, tile_size=16, eps2d=0.3, backgrounds=None, render_mode="RGB", absgrad=False,
                  sparse_grad=False, rasterize_mode="classic", camera_model="pinhole", distributed=False,
                  near_plane=None, far_plane=None):
    """See the module docstring. Unsupported options raise ValueError rather than render
    something else silently."""
    if packed:
        raise ValueError("packed=True is not supported (the reference calls packed=False)")
    if render_mode != "RGB" or backgrounds is not None or camera_model != "pinhole":
        raise ValueError("only render_mode='RGB', no backgrounds and the pinhole camera are supported")
    if absgrad or sparse_grad or distributed or rasterize_mode != "classic":
        raise ValueError("absgrad, sparse_grad, distributed and antialiased rasterization are not supported")
    if not math.isclose(eps2d, 0.3):
        raise ValueError("the 2D low-pass is the reference's 0.3")
    if tile_size not in (8, 16):
        raise ValueError("tile_size must be 8 or 16")
    n = means.shape[0]
    if colors.dim() == 2:  # post-activation RGB: degree-0 SH whose colour is max(0, rgb)
        colors = ((colors - 0.5) / SH_C0)[:, None, :]
        sh_degree = 0
    k = colors.shape[1]
    degree = int(round(math.sqrt(k))) - 1
    if (degree + 1) ** 2 != k or degree > 3:
        raise ValueError("colors must hold (d + 1)^2 SH coefficients, d <= 3")
    if sh_degree is None:
        sh_degree = degree
    cams = _cameras(viewmats, Ks, width, height)
    opts = RenderOptions(sh_degree=int(sh_degree), tile=(tile_size, tile_size))
    images, info = rasterize_views(means, scales, quats, opacities, colors, cams, opts, info=True, activated=True)
    C = len(cams)
    render_colors = torch.stack(images, 0)
    render_alphas = torch.stack(info["alphas"], 0)[..., None]
    radii = info["radii"][..., None].expand(C, n, 2)
    return render_colors, render_alphas, {"means2d": info["means2d"], "radii": radii, "width": int(width),
                                          "height": int(height), "n_cameras": C, "packed": False}


def viewmat_K(camera: Camera, device="cuda"):
    """The [4, 4] world-to-camera matrix and [3, 3] intrinsics of a Camera, as the training script
    builds them for each view."""
    vm = np.eye(4)
    vm[:3, :3] = np.asarray(camera.R, np.float64)
    vm[:3, 3] = np.asarray(camera.t, np.float64)
    K = np.array([[camera.fx, 0, camera.cx], [0, camera.fy, camera.cy], [0, 0, 1]], np.float64)
    return (torch.tensor(vm, dtype=torch.float32, device=device), torch.tensor(K, dtype=torch.float32, device=device))
