"""Edge cases of the forward path on the GPU, against the CPU oracle (key-exact mode).

Each case goes through check_frame (tests/test_gpu_forward.py): splat records, rects, keys,
per-tile lists and tile ranges bit for bit, images within 2/255. The cases are the ones a scene
or a caller can produce and the reference's code paths decide: non-finite and degenerate
parameters (the near test !(z >= 0.2) culls NaN depths; a non-finite centre or a radius of 2^24
pixels or more is culled; render.cpp:43-58), Gaussians on the near plane, a splat covering the
whole image, tiny and odd image sizes, a one-Gaussian and an empty scene, the SH degree option
above the scene's degree (clamped, sh.cpp:19), and a 3840x2160 frame with more than 2^15 tiles.
"""
import numpy as np
import pytest

import oracle
from aniso_amd import scenes
from aniso_amd.renderer import GpuRenderer, RenderOptions, RenderStats

from test_gpu_forward import check_frame

pytestmark = pytest.mark.gpu


def cam(w, h, fx=None):
    f = 0.8 * w if fx is None else fx
    return scenes.Camera(np.eye(3), np.zeros(3), f, f, w / 2.0, h / 2.0, w, h)


def scene_from(pos, scale, rot, opacity, sh):
    return oracle.ActivatedScene(pos, scale, rot, opacity, sh)


def base_scene(n, seed, degree=3):
    rng = np.random.default_rng(seed)
    z = rng.uniform(1.0, 10.0, n)
    pos = np.stack([rng.uniform(-0.6, 0.6, n) * z, rng.uniform(-0.4, 0.4, n) * z, z], 1)
    scale = np.exp(rng.normal(-3.0, 0.8, (n, 3)))
    q = rng.normal(0, 1, (n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    opacity = 1 / (1 + np.exp(-rng.normal(0, 1.5, n)))
    sh = rng.normal(0, 0.3, (n, (degree + 1) ** 2, 3))
    return pos, scale, q, opacity, sh


def test_degenerate_parameters(gpu):
    """NaN / Inf positions, zero and huge scales, opacity 0 and 1, Gaussians on and just inside
    the near plane, behind the camera, far away; among ordinary ones."""
    pos, scale, rot, opacity, sh = base_scene(4000, 1)
    pos[0] = [np.nan, 0, 2]
    pos[1] = [0, 0, np.nan]
    pos[2] = [np.inf, 0, 3]
    pos[3] = [0, 0, 0.2]          # on the near plane: kept (z >= 0.2)
    pos[4] = [0, 0, 0.19999999]   # just inside it: culled
    pos[5] = [0, 0, -1]           # behind
    pos[6] = [0, 0, 1e7]          # far
    scale[7] = 0.0                 # a point: only the 0.3 low-pass remains
    scale[8] = 1e-30
    scale[9] = [50.0, 50.0, 50.0]  # covers the whole image
    opacity[10] = 0.0
    opacity[11] = 1.0
    sc = scene_from(pos, scale, rot, opacity, sh)
    c = cam(320, 240)
    with GpuRenderer(sc, activated=True) as r:
        for tile in [(16, 16), (8, 8)]:
            check_frame(r, sc, c, tile=tile, label="degenerate")


@pytest.mark.parametrize("size", [(1, 1), (7, 5), (17, 1), (1, 33), (4096, 64)])
def test_tiny_and_odd_image_sizes(gpu, size):
    pos, scale, rot, opacity, sh = base_scene(3000, 2)
    sc = scene_from(pos, scale, rot, opacity, sh)
    with GpuRenderer(sc, activated=True) as r:
        check_frame(r, sc, cam(*size, fx=0.8 * max(size)), label=f"size {size}")


def test_one_gaussian_and_empty_scene(gpu):
    pos, scale, rot, opacity, sh = base_scene(1, 3)
    pos[0] = [0.0, 0.0, 2.0]
    scale[0] = 0.2
    sc = scene_from(pos, scale, rot, opacity, sh)
    with GpuRenderer(sc, activated=True) as r:
        img, st = check_frame(r, sc, cam(160, 120), label="one Gaussian")
        assert st.visible == 1 and img.any()
    empty = scene_from(np.zeros((0, 3)), np.zeros((0, 3)), np.zeros((0, 4)), np.zeros(0), np.zeros((0, 16, 3)))
    with GpuRenderer(empty, activated=True) as r:
        st = RenderStats()
        img = r.render(cam(160, 120), st)
        assert st.visible == 0 and st.tile_pairs == 0 and not img.any()


@pytest.mark.parametrize("scene_degree", [0, 1, 2])
def test_sh_option_above_scene_degree(gpu, scene_degree):
    pos, scale, rot, opacity, sh = base_scene(3000, 4, degree=scene_degree)
    sc = scene_from(pos, scale, rot, opacity, sh)
    with GpuRenderer(sc, activated=True) as r:
        a, _ = check_frame(r, sc, cam(320, 240), sh_degree=3, label=f"scene degree {scene_degree}, option 3")
        b = r.render(cam(320, 240), None, RenderOptions(sh_degree=scene_degree))
        assert np.array_equal(a, b)


def test_4k_frame_with_many_tiles(gpu):
    """3840x2160 at 8x8 tiles: 129,600 tiles (17-bit tile ids, three tile-sort passes)."""
    raw = scenes.synthetic_scene(200_000, 3840, 2160, seed=6)
    sc = oracle.activate(raw)
    c = scenes.pinhole_camera(3840, 2160)
    with GpuRenderer(sc, activated=True) as r:
        check_frame(r, sc, c, tile=(8, 8), label="4K")
