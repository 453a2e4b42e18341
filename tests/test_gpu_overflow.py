"""Pair-buffer overflow on the GPU: a frame whose pairs exceed the initial pair capacity
(max(8 n, 2^20) slots) must come out as if the buffers had been large enough.

1,200 Gaussians that each cover the whole 640x360 image touch 40 x 23 = 920 tiles: 1,104,000
pairs against a first capacity of 1,048,576. The synchronous render enqueues the rest of the
frame before it has the count, so the overflow is found with those kernels already queued: they
must skip, the buffers grow and the rest runs again (stats.retries = 1), with records, keys,
lists, ranges and image checked against the oracle. The asynchronous entry reports the overflow
from its wait (the reference reads the count back before it sorts, gpu_render.cu:324-329, and has
no such case), and the differentiable forward retries like the synchronous render (when its
pooled frame has not been grown by an earlier test in the process).
"""
import ctypes

import numpy as np
import pytest
import torch

import oracle
from aniso_amd import _lib, scenes
from aniso_amd.autograd import rasterize
from aniso_amd.renderer import GpuRenderer, RenderOptions, RenderStats, camera_c

from test_gpu_forward import check_frame

pytestmark = pytest.mark.gpu

W, H = 640, 360


def wide_scene(n=1200, seed=11):
    """Raw Gaussians 3-8 units in front of the camera with standard deviations of about 1.6
    units (hundreds of pixels): every one touches every tile."""
    rng = np.random.default_rng(seed)
    z = rng.uniform(3.0, 8.0, n)
    means = np.stack([rng.uniform(-0.2, 0.2, n) * z, rng.uniform(-0.1, 0.1, n) * z, z], 1).astype(np.float32)
    log_scales = rng.normal(0.5, 0.1, (n, 3)).astype(np.float32)
    q = rng.normal(0, 1, (n, 4))
    quats = (q / np.linalg.norm(q, axis=1, keepdims=True)).astype(np.float32)
    opacity_logits = rng.normal(-3.0, 0.5, n).astype(np.float32)
    sh = rng.normal(0, 0.3, (n, 16, 3)).astype(np.float32)
    return scenes.RawScene(means, log_scales, quats, opacity_logits, sh)


def test_sync_render_grows_and_matches_oracle(gpu):
    raw = wide_scene()
    act = oracle.activate(raw)
    cam = scenes.pinhole_camera(W, H)
    with GpuRenderer(act, activated=True) as r:
        img, st = check_frame(r, act, cam, label="overflow")
        print(f"  pairs {st.tile_pairs}, capacity after {st.pair_capacity}, retries {st.retries}", flush=True)
        assert st.tile_pairs > (1 << 20) and st.retries == 1 and st.pair_capacity >= st.tile_pairs
        st2 = RenderStats()
        img2 = r.render(cam, st2)
        assert st2.retries == 0 and np.array_equal(img, img2)


def test_async_render_reports_overflow_then_renders(gpu):
    raw = wide_scene()
    cam = scenes.pinhole_camera(W, H)
    L = _lib.load()
    dev = torch.empty(W * H * 3, dtype=torch.uint8, device="cuda")
    c_cam, c_opt = camera_c(cam), RenderOptions().c()
    with GpuRenderer(raw) as ref_r:
        want = ref_r.render(cam)
    with GpuRenderer(raw) as r:
        _lib.check(L.ans_render_async(r._h, c_cam, c_opt, dev.data_ptr(), None))
        rc = L.ans_render_wait(r._h, None)
        assert rc == 3, rc  # ANS_ERR_CAPACITY
        _lib.check(L.ans_render_async(r._h, c_cam, c_opt, dev.data_ptr(), None))
        st = _lib.Stats_()
        _lib.check(L.ans_render_wait(r._h, ctypes.byref(st)))
        torch.cuda.synchronize()
        assert st.tile_pairs > (1 << 20)
        assert np.array_equal(dev.cpu().numpy().reshape(H, W, 3), want)


def test_async_overflow_grows_to_the_largest_frame(gpu):
    """Frames [big, empty, empty, empty] on the three asynchronous streams: the fourth frame runs
    on the first frame's stream after it, so the buffers must grow to the overflowing frame's
    pair count, not to the count of the frame that ran last on that stream. The wait reports
    ANS_ERR_CAPACITY once; the same four frames rendered again all succeed and equal their
    synchronous renders."""
    raw = wide_scene()
    big = scenes.pinhole_camera(W, H)
    behind = scenes.pinhole_camera(W, H)
    behind.R = np.diag([-1.0, 1.0, -1.0])
    cams = [big, behind, behind, behind]
    L = _lib.load()
    with GpuRenderer(raw) as ref_r:
        want = [ref_r.render(c) for c in cams]
    assert want[0].any() and not any(w.any() for w in want[1:])
    outs = [torch.full((W * H * 3,), 7, dtype=torch.uint8, device="cuda") for _ in cams]
    c_opt = RenderOptions().c()
    with GpuRenderer(raw) as r:
        for c, o in zip(cams, outs):
            _lib.check(L.ans_render_async(r._h, camera_c(c), c_opt, o.data_ptr(), None))
        rc = L.ans_render_wait(r._h, None)
        assert rc == 3, rc  # ANS_ERR_CAPACITY
        for c, o in zip(cams, outs):
            _lib.check(L.ans_render_async(r._h, camera_c(c), c_opt, o.data_ptr(), None))
        st = _lib.Stats_()
        _lib.check(L.ans_render_wait(r._h, ctypes.byref(st)))
        torch.cuda.synchronize()
        assert st.pair_capacity > (1 << 20)
        for k, (o, w) in enumerate(zip(outs, want)):
            assert np.array_equal(o.cpu().numpy().reshape(H, W, 3), w), f"frame {k}"


def test_differentiable_forward_retries(gpu):
    raw = wide_scene()
    cam = scenes.pinhole_camera(W, H)
    t = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float32, device="cuda", requires_grad=True)
    params = [t(raw.means), t(raw.log_scales), t(raw.quats), t(raw.opacity_logits), t(raw.sh)]
    st = RenderStats()
    img = rasterize(*params, cam, None, st)
    # The differentiable handle reuses released frames, which keep the capacity earlier tests in
    # the process gave them: the retry happens when this test runs first.
    print(f"pairs {st.tile_pairs}, retries {st.retries}", flush=True)
    assert st.tile_pairs > (1 << 20) and st.retries in (0, 1)
    img.sum().backward()
    assert all(torch.isfinite(p.grad).all() for p in params)
    grads = [p.grad.clone() for p in params]
    for p in params:
        p.grad = None
    st2 = RenderStats()
    img2 = rasterize(*params, cam, None, st2)
    img2.sum().backward()
    assert torch.equal(img, img2)
    for a, p in zip(grads, params):
        assert torch.equal(a, p.grad)
    # The render API activates the raw parameters on upload, the differentiable forward in its
    # preprocess: the images agree to rounding.
    with GpuRenderer(raw) as r:
        _, ref = r.render_float(cam)
    err = np.abs(img.detach().cpu().numpy() - ref).max()
    print(f"max |differentiable - render API| {err:.2e}", flush=True)
    assert err < 1e-3
