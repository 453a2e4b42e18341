"""The forward that allocates nothing (ans_diff_forward_noalloc), through the C ABI, against the
differentiable forward it is a variant of (ans_diff_forward, through rasterize): the float images
must be bit-identical, on two calls in a row, the second with a smaller scene and image, then on
the first again. Also: the Gaussian limit is refused before any GPU work, and a frame whose pairs
do not fit the fixed scratch returns ANS_ERR_CAPACITY."""
import ctypes

import numpy as np
import pytest
import torch

from aniso_amd import _lib, scenes
from aniso_amd.autograd import _params, rasterize
from aniso_amd.renderer import RenderOptions, camera_c

pytestmark = pytest.mark.gpu


def tensors(raw):
    t = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float32, device="cuda")
    return [t(raw.means), t(raw.log_scales), t(raw.quats), t(raw.opacity_logits), t(raw.sh)]


def noalloc(params, cam, opts=None):
    L = _lib.load()
    img = torch.full((cam.height, cam.width, 3), -7.0, dtype=torch.float32, device="cuda")
    st = _lib.Stats_()
    rc = L.ans_diff_forward_noalloc(ctypes.byref(_params(*params)), ctypes.byref(camera_c(cam)),
                                    ctypes.byref((opts or RenderOptions()).c()), img.data_ptr(),
                                    ctypes.c_void_p(torch.cuda.current_stream().cuda_stream), ctypes.byref(st))
    return rc, img, st


def check(raw, cam, label):
    params = tensors(raw)
    print(f"case {label}: {raw.n} Gaussians {cam.width}x{cam.height}", flush=True)
    rc, img, st = noalloc(params, cam)
    assert rc == 0, _lib.load().ans_last_error()
    with torch.no_grad():
        ref = rasterize(*params, cam)
    torch.cuda.synchronize()
    print(f"  pairs {st.tile_pairs} of {st.pair_capacity}", flush=True)
    assert torch.equal(img.view(torch.int32), ref.view(torch.int32)), f"{label}: images differ"


def test_noalloc_matches_diff_forward_twice(gpu):
    big = scenes.synthetic_scene(200_000, 640, 360, seed=1)
    small = scenes.synthetic_scene(30_000, 320, 180, seed=2)
    cam_big = scenes.pose_camera([0.1, -0.05, 0.0], [0.05, 0.0, 1.0], [0, -1, 0], 640, 360, 60.0)
    cam_small = scenes.pose_camera([0.0, 0.0, 0.1], [0.0, 0.02, 1.0], [0, -1, 0], 320, 180, 60.0)
    check(big, cam_big, "first")
    check(small, cam_small, "second, smaller")
    check(big, cam_big, "first again")


def test_noalloc_limits(gpu):
    L = _lib.load()
    n = (1 << 19) + 1
    params = [torch.zeros((n, 3), device="cuda"), torch.zeros((n, 3), device="cuda"),
              torch.zeros((n, 4), device="cuda"), torch.zeros(n, device="cuda"), torch.zeros((n, 1, 3), device="cuda")]
    cam = scenes.pinhole_camera(64, 64)
    img = torch.empty((64, 64, 3), device="cuda")
    rc = L.ans_diff_forward_noalloc(ctypes.byref(_params(*params)), ctypes.byref(camera_c(cam)), None,
                                    img.data_ptr(), None, None)
    assert rc == 1 and b"524288" in L.ans_last_error()
    # 6,000 Gaussians that each cover all 920 tiles of a 640x360 image: 5.5M pairs, more than the
    # scratch holds.
    rng = np.random.default_rng(3)
    k = 6000
    z = rng.uniform(3.0, 8.0, k)
    means = np.stack([rng.uniform(-0.2, 0.2, k) * z, rng.uniform(-0.1, 0.1, k) * z, z], 1)
    wide = scenes.RawScene(means.astype(np.float32), rng.normal(0.5, 0.1, (k, 3)).astype(np.float32),
                           np.tile(np.float32([1, 0, 0, 0]), (k, 1)), np.full(k, -3.0, np.float32),
                           rng.normal(0, 0.3, (k, 16, 3)).astype(np.float32))
    rc, img, st = noalloc(tensors(wide), scenes.pinhole_camera(640, 360))
    print(f"wide: pairs {st.tile_pairs} capacity {st.pair_capacity} rc {rc}", flush=True)
    assert rc == 3 and st.tile_pairs > st.pair_capacity
    assert bool((img == -7.0).all())  # left unwritten
    # The entry works again afterwards.
    small = scenes.synthetic_scene(30_000, 320, 180, seed=4)
    check(small, scenes.pinhole_camera(320, 180), "after an overflow")
