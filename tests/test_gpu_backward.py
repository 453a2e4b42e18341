"""Backward parity on the GPU: gradients of the HIP backward kernels (through the
torch.autograd.Function and the C ABI) against the FP64 CPU reference gradient evaluated on the
GPU's own per-tile lists, and against FP64 central finite differences. Tolerance: 1e-3
relative error per parameter group (norm of the difference over norm of the reference); the
median and 99th percentile of the per-element relative error are printed."""
import numpy as np
import pytest
import torch

import oracle
from aniso_amd import scenes
from aniso_amd.autograd import frame_lists, rasterize
from aniso_amd.renderer import RenderOptions, RenderStats

pytestmark = pytest.mark.gpu

TOL = 1e-3


def tensors(raw, device="cuda"):
    t = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float32, device=device, requires_grad=True)
    return [t(raw.means), t(raw.log_scales), t(raw.quats), t(raw.opacity_logits), t(raw.sh)]


def gpu_grads(params, cam, dLdC, options=None):
    for p in params:
        p.grad = None
    img = rasterize(*params, cam, options)
    w = torch.tensor(dLdC, dtype=torch.float32, device="cuda")
    (img * w).sum().backward()
    g = np.concatenate([params[0].grad.cpu().numpy(), params[1].grad.cpu().numpy(), params[2].grad.cpu().numpy(),
                        params[3].grad.cpu().numpy()[:, None],
                        params[4].grad.cpu().numpy().reshape(len(params[0]), -1)], 1)
    return img.detach().cpu().numpy(), g.astype(np.float64)


def compare(name_prefix, g, ref, rows=None):
    worst = 0.0
    for name, sl in oracle.PARAM_GROUPS.items():
        a, b = g[:, sl], ref[:, sl]
        if rows is not None:
            a, b = a[rows], b[rows]
        err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        el = np.abs(a - b) / np.maximum(np.abs(b), 1e-3 * np.abs(b).max() + 1e-30)
        print(f"  {name_prefix} {name:15s} rel {err:.2e}  median {np.median(el):.1e}  p99 {np.quantile(el, 0.99):.1e}",
              flush=True)
        worst = max(worst, err)
    return worst


def l1_case(n, w, h, seed, scale_boost=0.0):
    raw = scenes.synthetic_scene(n, w, h, seed=seed)
    raw.log_scales += np.float32(scale_boost)
    cam = scenes.pose_camera([0.05, -0.05, -0.5], [0.05, 0.02, 1.0], [0, -1, 0], w, h, 60.0)
    return raw, cam


@pytest.mark.parametrize("tile", [(16, 16), (8, 8), (32, 16)])
def test_gradients_match_fp64_reference_small(gpu, tile):
    raw, cam = l1_case(20000, 320, 180, seed=3, scale_boost=1.5)
    print(f"case small-20k tile={tile}", flush=True)
    params = tensors(raw)
    opts = RenderOptions(tile=tile)
    start, values, _ = frame_lists(*[p.detach() for p in params], cam, opts)
    rng = np.random.default_rng(0)
    img0, _, _ = oracle.grad_run(raw, cam, np.zeros((cam.height, cam.width, 3)), tile=tile, lists=(start, values))
    dLdC = np.sign(img0 - (img0 + rng.normal(0, 0.05, img0.shape)))
    img, g = gpu_grads(params, cam, dLdC, opts)
    ref_img, ref, _ = oracle.grad_run(raw, cam, dLdC, tile=tile, lists=(start, values))
    d = np.abs(img - ref_img)
    print(f"  image |GPU FP32 - FP64|: max {d.max():.2e} mean {d.mean():.2e}", flush=True)
    assert d.max() < 1.0 / 255.0 and d.mean() < 1e-5
    assert compare("vs FP64 reference", g, ref) < TOL


def test_gradients_match_central_differences(gpu):
    raw, cam = l1_case(20000, 320, 180, seed=5, scale_boost=1.5)
    params = tensors(raw)
    rng = np.random.default_rng(1)
    img0, _, _ = oracle.grad_run(raw, cam, np.zeros((cam.height, cam.width, 3)))
    dLdC = np.sign(img0 - (img0 + rng.normal(0, 0.05, img0.shape)))
    _, g = gpu_grads(params, cam, dLdC)
    idx = rng.choice(np.flatnonzero(np.abs(g).sum(1) > 0), 150, replace=False)
    fd, flags = oracle.grad_fd(raw, cam, dLdC, idx)
    ok = flags == 0
    print(f"  FD samples kept {ok.mean():.3f}", flush=True)
    worst = 0.0
    for name, sl in oracle.PARAM_GROUPS.items():
        a, b = g[idx][:, sl][ok[:, sl]], fd[:, sl][ok[:, sl]]
        err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        print(f"  vs FP64 central differences {name:15s} rel {err:.2e}", flush=True)
        worst = max(worst, err)
    assert worst < TOL


def test_zero_image_gradient_gives_zero_gradients(gpu):
    raw, cam = l1_case(5000, 160, 90, seed=8, scale_boost=1.0)
    params = tensors(raw)
    _, g = gpu_grads(params, cam, np.zeros((90, 160, 3)))
    assert not g.any()


def test_culled_gaussians_get_zero_gradient(gpu):
    raw, cam = l1_case(5000, 160, 90, seed=9, scale_boost=1.0)
    raw.means[:100, 2] = -10.0
    params = tensors(raw)
    _, g = gpu_grads(params, cam, np.ones((90, 160, 3)))
    assert not g[:100].any() and g[100:].any()


def test_backward_is_deterministic(gpu):
    raw, cam = l1_case(20000, 320, 180, seed=4, scale_boost=1.5)
    params = tensors(raw)
    d = np.random.default_rng(2).normal(size=(180, 320, 3))
    _, a = gpu_grads(params, cam, d)
    _, b = gpu_grads(params, cam, d)
    assert np.array_equal(a, b)


def high_opacity_case(seed=13):
    """Opacity logits shifted by +4 (opacity about 0.98 on average, so most alphas near a splat's
    centre reach the 0.99 clamp) plus 40 splats made about ten times larger, so that many pixels
    stop on T < 1e-4 after a few clamped splats: the regime of trained scenes, where the
    transmittance the adjoint recovers by division through (1 - alpha) is most sensitive."""
    raw, cam = l1_case(20000, 320, 180, seed=seed, scale_boost=1.5)
    raw.opacity_logits += np.float32(4.0)
    rng = np.random.default_rng(seed)
    big = rng.choice(raw.n, 40, replace=False)
    raw.log_scales[big] += np.float32(2.3)
    return raw, cam


def test_high_opacity_gradients_match_fp64(gpu):
    raw, cam = high_opacity_case()
    params = tensors(raw)
    start, values, _ = frame_lists(*[p.detach() for p in params], cam)
    img0, _, _ = oracle.grad_run(raw, cam, np.zeros((cam.height, cam.width, 3)), lists=(start, values))
    rng = np.random.default_rng(4)
    dLdC = np.sign(img0 - (img0 + rng.normal(0, 0.05, img0.shape)))
    img, g = gpu_grads(params, cam, dLdC)
    ref_img, ref, _ = oracle.grad_run(raw, cam, dLdC, lists=(start, values))
    op = 1.0 / (1.0 + np.exp(-raw.opacity_logits.astype(np.float64)))
    print(f"case high-opacity: mean opacity {op.mean():.3f}, image |GPU - FP64| max "
          f"{np.abs(img - ref_img).max():.2e}", flush=True)
    assert np.abs(img - ref_img).max() < 1.0 / 255.0
    assert compare("vs FP64 reference", g, ref) < TOL
    idx = rng.choice(np.flatnonzero(np.abs(ref).sum(1) > 0), 150, replace=False)
    fd, flags = oracle.grad_fd(raw, cam, dLdC, idx)
    ok = flags == 0
    print(f"  FD samples kept {ok.mean():.3f}", flush=True)
    worst = 0.0
    for name, sl in oracle.PARAM_GROUPS.items():
        a, b = g[idx][:, sl][ok[:, sl]], fd[:, sl][ok[:, sl]]
        if b.size == 0:
            continue
        err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        print(f"  vs FP64 central differences {name:15s} rel {err:.2e}", flush=True)
        worst = max(worst, err)
    assert worst < TOL


@pytest.fixture(scope="module")
def c4_case():
    """741,883 Gaussians (SH3) at 1280x720; L1 against a target rendered after perturbing the
    means by N(0, 0.01^2) and the SH DC by N(0, 0.05^2)."""
    raw = scenes.synthetic_scene(741_883, 1280, 720)
    cam = scenes.pinhole_camera(1280, 720)
    rng = np.random.default_rng(0)
    tgt = scenes.RawScene(raw.means + rng.normal(0, 0.01, raw.means.shape).astype(np.float32), raw.log_scales,
                          raw.quats, raw.opacity_logits, raw.sh.copy())
    tgt.sh[:, 0, :] += rng.normal(0, 0.05, tgt.sh[:, 0, :].shape).astype(np.float32)
    with torch.no_grad():
        target = rasterize(*[p.detach() for p in tensors(tgt)], cam)
    return raw, cam, target


def test_741883_l1_gradients_vs_fp64_reference(c4_case):
    raw, cam, target = c4_case
    params = tensors(raw)
    start, values, _ = frame_lists(*[p.detach() for p in params], cam)
    for p in params:
        p.grad = None
    st = RenderStats()
    img = rasterize(*params, cam, None, st)
    loss = (img - target).abs().mean()
    loss.backward()
    dLdC = (torch.sign(img.detach() - target) / img.numel()).cpu().numpy().astype(np.float64)
    g = np.concatenate([params[0].grad.cpu().numpy(), params[1].grad.cpu().numpy(), params[2].grad.cpu().numpy(),
                        params[3].grad.cpu().numpy()[:, None], params[4].grad.cpu().numpy().reshape(raw.n, -1)],
                       1).astype(np.float64)
    _, ref, vis = oracle.grad_run(raw, cam, dLdC, lists=(start, values))
    print(f"case 741883 visible {vis} pairs {st.tile_pairs}", flush=True)
    assert compare("vs FP64 reference", g, ref) < TOL
    # Central differences on 1,000 random Gaussians.
    rng = np.random.default_rng(7)
    idx = rng.choice(np.flatnonzero(np.abs(ref).sum(1) > 0), 1000, replace=False)
    fd, flags = oracle.grad_fd(raw, cam, dLdC, idx)
    ok = flags == 0
    print(f"  FD samples kept {ok.mean():.3f}", flush=True)
    worst = 0.0
    for name, sl in oracle.PARAM_GROUPS.items():
        a, b = g[idx][:, sl][ok[:, sl]], fd[:, sl][ok[:, sl]]
        err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        print(f"  vs FP64 central differences {name:15s} rel {err:.2e}", flush=True)
        worst = max(worst, err)
    assert worst < TOL


def test_info_radii_and_screen_gradients(gpu):
    """rasterize(info=True): radii equal the oracle's (key-exact mode on the host-activated scene;
    device and host activation may differ in a last bit, so a handful of radii may differ by one),
    culled Gaussians have radius 0, and means2d.grad equals dL/d(u, v) of the splat-record gradient
    from the same kernels, and the FP64 reference's within the gradient tolerance."""
    raw, cam = l1_case(20000, 320, 180, seed=11, scale_boost=1.0)
    rng = np.random.default_rng(5)
    dLdC = rng.normal(size=(cam.height, cam.width, 3)).astype(np.float32)
    params = tensors(raw)
    img, info = rasterize(*params, cam, info=True)
    (img * torch.tensor(dLdC, device="cuda")).sum().backward()
    radii = info["radii"].cpu().numpy()
    g2d = info["means2d"].grad.cpu().numpy()
    with oracle.Frame(oracle.activate(raw), cam, oracle.KEYEXACT, tile=(16, 16)) as f:
        _, _, radius, alive = f.splats()
    alive = alive.astype(bool)
    assert np.all(radii[~alive] == 0)
    diff = radii[alive] != radius[alive]
    print(f"radii: {alive.sum()} visible, {diff.sum()} differ", flush=True)
    assert diff.sum() <= max(2, alive.sum() // 10000)
    assert np.all(np.abs(radii[alive].astype(int) - radius[alive].astype(int)) <= 1)
    # Same kernels, same frame: the splat-record gradient's first two columns.
    from aniso_amd.autograd import forward_backward
    with torch.no_grad():
        _, _, splat = forward_backward(*[p.detach() for p in params], cam,
                                       torch.tensor(dLdC, device="cuda"))
    assert np.array_equal(g2d, splat[:, :2])
    # FP64 reference on the GPU's lists.
    start, lst, _ = frame_lists(*[p.detach() for p in params], cam)
    _, _, _, sg = oracle.grad_run(raw, cam, dLdC, lists=(start, lst), splat_grad=True)
    err = np.linalg.norm(g2d - sg[:, :2]) / np.linalg.norm(sg[:, :2])
    print(f"means2d grad rel err vs FP64: {err:.2e}", flush=True)
    assert err < TOL
