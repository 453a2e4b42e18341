"""Camera-path streaming on the GPU (ans_render_path), against single-frame renders.

Contract: every frame handed to the sink equals ans_render (host RGB8) of that camera with that
frame's splat scale, bit for bit; frames arrive in path order, each exactly once; a sink that
stops or raises ends the call with an error after the GPU work drained, and the handle stays
usable. The reference's `aniso path` (src/main.cpp:206-250) renders the same frames one
GpuRenderer::render at a time and writes each to stdout.
"""
import dataclasses
import io

import numpy as np
import pytest

from aniso_amd import scenes
from aniso_amd.renderer import GpuRenderer, RenderOptions

pytestmark = pytest.mark.gpu


def path_cameras(seed, count, sizes=((320, 180),)):
    """A smooth walk through the synthetic scene, like a flythrough path file."""
    rng = np.random.default_rng(seed)
    start = rng.uniform(-0.3, 0.3, 3)
    cams = []
    for k in range(count):
        w, h = sizes[k % len(sizes)]
        yaw = np.radians(-15 + 30 * k / max(count - 1, 1))
        pos = start + np.array([0.02 * k, 0.0, 0.01 * k])
        cams.append(scenes.pose_camera(pos, (np.sin(yaw), 0.0, np.cos(yaw)), (0.0, -1.0, 0.0), w, h, 60.0))
    return cams


def reveal_scales(count):
    """A reveal: runs of equal scales, a shrink to 0.3 and back (the path file's tenth column)."""
    t = np.linspace(0.0, 1.0, count)
    s = 1.0 - 0.7 * np.sin(np.pi * t) ** 2
    s[: count // 4] = 1.0
    return s.astype(np.float32)


class Collect:
    def __init__(self):
        self.frames = []

    def __call__(self, index, rgb):
        self.frames.append((index, rgb.copy()))


def check_against_single(r, cams, scales, opts, frames):
    assert [i for i, _ in frames] == list(range(len(cams))), "frames out of order or missing"
    for (i, img), cam in zip(frames, cams):
        o = opts if scales is None else dataclasses.replace(opts, splat_scale=float(scales[i]))
        ref = r.render(cam, None, o)
        assert img.shape == ref.shape
        assert np.array_equal(img, ref), f"frame {i} differs from its single-frame render"


@pytest.fixture(scope="module")
def scene():
    return scenes.synthetic_scene(120_000, 640, 360, seed=5)


@pytest.fixture(scope="module")
def r(gpu, scene):
    with GpuRenderer(scene) as rr:
        yield rr


@pytest.mark.parametrize("chunk", [1, 8, 16])
def test_path_frames_equal_single_renders(r, chunk):
    cams = path_cameras(1, 21)
    scales = reveal_scales(len(cams))
    print(f"case path chunk={chunk}: {len(cams)} frames", flush=True)
    sink = Collect()
    st = r.render_path(cams, sink, RenderOptions(), chunk=chunk, splat_scales=scales)
    assert st.frames == len(cams)
    assert st.chunks == -(-len(cams) // chunk)
    check_against_single(r, cams, scales, RenderOptions(), sink.frames)


def test_path_mixed_sizes_and_options(r):
    cams = path_cameras(2, 11, sizes=((320, 180), (333, 217), (97, 61)))
    opts = RenderOptions(sh_degree=1, tile=(8, 8), batch_depth=64, splat_scale=0.8)
    print("case path mixed sizes", flush=True)
    sink = Collect()
    r.render_path(cams, sink, opts, chunk=4)
    check_against_single(r, cams, None, opts, sink.frames)


def test_path_file_sink_writes_raw_rgb(r):
    """A file object receives the raw RGB stream `aniso path` writes to stdout."""
    cams = path_cameras(3, 6)
    out = io.BytesIO()
    print("case path file sink", flush=True)
    r.render_path(cams, out, chunk=4)
    expect = b"".join(r.render(c).tobytes() for c in cams)
    assert out.getvalue() == expect


def test_path_sink_stop_and_error(r):
    cams = path_cameras(4, 10)
    seen = []

    def stop_at_3(i, rgb):
        seen.append(i)
        return False if i == 3 else None

    print("case path sink stop", flush=True)
    with pytest.raises(ValueError, match="frame 3"):
        r.render_path(cams, stop_at_3, chunk=4)
    assert seen == [0, 1, 2, 3]

    def boom(i, rgb):
        if i == 5:
            raise KeyError("sink failed")

    with pytest.raises(KeyError):
        r.render_path(cams, boom, chunk=4)
    sink = Collect()  # the handle still works
    r.render_path(cams, sink, chunk=4)
    check_against_single(r, cams, None, RenderOptions(), sink.frames)


def test_path_empty_and_bad_arguments(r):
    st = r.render_path([], Collect())
    assert st.frames == 0 and st.chunks == 0
    cams = path_cameras(5, 3)
    with pytest.raises(ValueError):
        r.render_path(cams, Collect(), chunk=0)
    with pytest.raises(ValueError):
        r.render_path(cams, Collect(), splat_scales=[1.0, -1.0, 1.0])


def test_path_pair_arena_grows_mid_path(gpu):
    """Large splats: a view's pairs (about 2.1M) exceed the first arena (max(8 n, 2^20) per view);
    the chunk is rendered again after the arena grows, and every frame is still exact."""
    sc = scenes.synthetic_scene(30_000, 640, 360, seed=9)
    sc.log_scales[:] += 2.0
    cams = path_cameras(6, 7, sizes=((640, 360),))
    print("case path arena growth", flush=True)
    with GpuRenderer(sc) as rr:
        sink = Collect()
        st = rr.render_path(cams, sink, RenderOptions(tile=(8, 8)), chunk=3)
        assert st.chunks > 3, "expected a re-render after arena growth"
        check_against_single(rr, cams, None, RenderOptions(tile=(8, 8)), sink.frames)
