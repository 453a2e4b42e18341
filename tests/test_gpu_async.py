"""Asynchronous frames on the GPU: ans_render_async pipelines consecutive frames over two streams
with their own frame buffers (one frame's sorts overlap the previous frame's blend), and every
frame must still equal its synchronous render, finish in enqueue order, and be reported by
ans_render_wait as the last frame.

The reference renders one frame per render() call (gpu_render.cu:294-375), so its results are
the ones a frame loop gets: each image here is compared bit for bit with the synchronous render
of the same camera, which the forward tests check against the oracle.
"""
import ctypes

import numpy as np
import pytest
import torch

from aniso_amd import _lib, scenes
from aniso_amd.renderer import GpuRenderer, RenderOptions, RenderStats, camera_c

pytestmark = pytest.mark.gpu

W, H = 640, 360


def cameras(k):
    return [scenes.pose_camera([0.05 * i, -0.03 * i, 0.1 * i], [0.02 * i, 0.01, 1.0], [0, -1, 0], W, H, 60.0)
            for i in range(k)]


def test_async_frames_equal_sync_renders(gpu):
    raw = scenes.synthetic_scene(200_000, W, H, seed=5)
    cams = cameras(7)
    L = _lib.load()
    with GpuRenderer(raw) as r:
        want = []
        for c in cams:
            st = RenderStats()
            want.append((r.render(c, st), st.visible, st.tile_pairs))
        outs = [torch.empty(W * H * 3, dtype=torch.uint8, device="cuda") for _ in cams]
        opt = RenderOptions().c()
        for c, o in zip(cams, outs):
            _lib.check(L.ans_render_async(r._h, camera_c(c), opt, o.data_ptr(), None))
        st = _lib.Stats_()
        _lib.check(L.ans_render_wait(r._h, ctypes.byref(st)))
        torch.cuda.synchronize()
        for i, (o, (img, _, _)) in enumerate(zip(outs, want)):
            assert np.array_equal(o.cpu().numpy().reshape(H, W, 3), img), f"frame {i}"
        assert st.visible == want[-1][1] and st.tile_pairs == want[-1][2]
        assert st.project_ms > 0 and st.sort_ms > 0 and st.blend_ms > 0
        # The last frame's intermediates are the ones dumped.
        _, vis, pairs, _ = r.frame_counts()
        assert vis == want[-1][1] and pairs == want[-1][2]


def test_async_frames_into_one_output_finish_in_order(gpu):
    """Frames of alternating cameras into the same device image: the image after the wait is the
    last frame's (blends run in enqueue order across the two streams)."""
    raw = scenes.synthetic_scene(200_000, W, H, seed=6)
    cams = cameras(2)
    L = _lib.load()
    with GpuRenderer(raw) as r:
        want = [r.render(c) for c in cams]
        out = torch.empty(W * H * 3, dtype=torch.uint8, device="cuda")
        opt = RenderOptions().c()
        for k in range(9):
            _lib.check(L.ans_render_async(r._h, camera_c(cams[k % 2]), opt, out.data_ptr(), None))
            if k in (4, 8):
                _lib.check(L.ans_render_wait(r._h, None))
                torch.cuda.synchronize()
                assert np.array_equal(out.cpu().numpy().reshape(H, W, 3), want[k % 2]), f"after frame {k}"


def test_sync_and_async_frames_interleaved(gpu):
    raw = scenes.synthetic_scene(100_000, W, H, seed=7)
    cams = cameras(4)
    L = _lib.load()
    with GpuRenderer(raw) as r:
        want = [r.render(c) for c in cams]
        out = torch.empty(W * H * 3, dtype=torch.uint8, device="cuda")
        opt = RenderOptions().c()
        for k, c in enumerate(cams):
            _lib.check(L.ans_render_async(r._h, camera_c(c), opt, out.data_ptr(), None))
            img = r.render(cams[(k + 1) % len(cams)])
            assert np.array_equal(img, want[(k + 1) % len(cams)])
            _lib.check(L.ans_render_wait(r._h, None))
            torch.cuda.synchronize()
            assert np.array_equal(out.cpu().numpy().reshape(H, W, 3), want[k]), f"async frame {k}"


def test_async_full_size_clustered_1080p(gpu):
    """2.2M clustered Gaussians at 1920x1080 (11.5M pairs per frame): three poses pipelined over
    the three streams equal their synchronous renders."""
    w, h = 1920, 1080
    raw = scenes.synthetic_scene(2_200_000, w, h, clustered=True)
    cams = [scenes.pose_camera([0.04 * i, -0.02 * i, 0.05 * i], [0.01 * i, 0.0, 1.0], [0, -1, 0], w, h, 60.0)
            for i in range(3)]
    L = _lib.load()
    with GpuRenderer(raw) as r:
        want = [r.render(c) for c in cams]
        outs = [torch.empty(w * h * 3, dtype=torch.uint8, device="cuda") for _ in cams]
        opt = RenderOptions().c()
        for _ in range(2):
            for c, o in zip(cams, outs):
                _lib.check(L.ans_render_async(r._h, camera_c(c), opt, o.data_ptr(), None))
        _lib.check(L.ans_render_wait(r._h, None))
        torch.cuda.synchronize()
        for i, (o, img) in enumerate(zip(outs, want)):
            assert np.array_equal(o.cpu().numpy().reshape(h, w, 3), img), f"frame {i}"


def test_destroy_with_frames_in_flight(gpu):
    """Destroying a handle right after enqueueing frames (no wait) drains them first; a new
    handle then renders correctly."""
    raw = scenes.synthetic_scene(200_000, W, H, seed=8)
    cam = cameras(1)[0]
    L = _lib.load()
    out = torch.empty(W * H * 3, dtype=torch.uint8, device="cuda")
    opt = RenderOptions().c()
    for _ in range(3):
        r = GpuRenderer(raw)
        for _ in range(5):
            _lib.check(L.ans_render_async(r._h, camera_c(cam), opt, out.data_ptr(), None))
        r.close()
    with GpuRenderer(raw) as r:
        want = r.render(cam)
        _lib.check(L.ans_render_async(r._h, camera_c(cam), opt, out.data_ptr(), None))
        _lib.check(L.ans_render_wait(r._h, None))
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy().reshape(H, W, 3), want)
