"""PLY ingest and export on the GPU, against the reference's own loader.

Contract: a renderer made from a PLY file (ans_renderer_load_ply, streamed; or
ans_renderer_create_ply, from bytes) holds exactly the arrays the reference loader
(aniso::loadPly, compiled from src/ply_loader.cpp into oracle/_ref) produces for that file: the
same float32 bits for positions, activated scales, opacities, unit quaternions and SH, including
extreme raw values; degree-4 files keep degrees 0..3. Renders of it equal renders of the same
scene activated on the host. The writer (ans_ply_write) produces the training pipeline's file
layout (pipeline/train_gsplat.py:109-128) byte for byte, and its output loads back through the
reference loader.
"""
import numpy as np
import pytest
import torch

import oracle
from aniso_amd import scenes
from aniso_amd.autograd import save_ply
from aniso_amd.renderer import GpuRenderer, RenderOptions

from test_ply import ply_file, reference_test_scene, standard_props, valid_variants

pytestmark = pytest.mark.gpu

needs_ref = pytest.mark.skipif(not oracle.ref_available(), reason="needs the compiled reference loader")


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def check_equal_to_reference(r, path):
    ref, _ = oracle.ref_load_ply(path)
    got = r.scene_arrays()
    assert r.n == ref["pos"].shape[0]
    k = (min(ref["sh_degree"], 3) + 1) ** 2
    assert r.sh_degree == min(ref["sh_degree"], 3)
    for key in ("pos", "scale", "rot", "opacity"):
        assert np.array_equal(bits(got[key]), bits(ref[key])), f"{key} differs from the reference loader"
    assert np.array_equal(bits(got["sh"]), bits(ref["sh"][:, :k, :])), "sh differs from the reference loader"


def extreme_scene(n, degree, seed):
    """Raw values over the whole range the activation sees: log-scales and logits from -120 to
    120 (overflow, underflow, subnormal results), zero and tiny quaternions."""
    rng = np.random.default_rng(seed)
    sc = scenes.synthetic_scene(n, 320, 240, seed=seed)
    sc.sh = np.ascontiguousarray(sc.sh[:, : (degree + 1) ** 2, :])
    special = np.array([-120, -104, -103.9, -100, -90, -88, -87.5, -30, -1e-30, 0, 1e-30, 30, 88, 88.7, 89, 120],
                       np.float32)
    sc.log_scales[: special.size] = special[:, None]
    sc.opacity_logits[: special.size] = special
    sc.log_scales[special.size:] = rng.uniform(-120, 120, sc.log_scales[special.size:].shape)
    sc.quats[0] = 0.0
    sc.quats[1] = [1e-30, 0, 0, 0]
    sc.quats[2] = [3e19, 4e19, 0, 0]
    return sc


@needs_ref
@pytest.mark.parametrize("degree", [0, 1, 2, 3])
def test_load_matches_reference_loader(gpu, tmp_path, degree):
    sc = extreme_scene(5000, degree, seed=degree)
    path = tmp_path / f"deg{degree}.ply"
    scenes.write_ply(path, sc)
    print(f"case load degree {degree}", flush=True)
    with GpuRenderer.from_ply(str(path)) as r:
        check_equal_to_reference(r, path)
    with GpuRenderer.from_ply(path.read_bytes()) as r:
        check_equal_to_reference(r, path)


@needs_ref
@pytest.mark.parametrize("degree", [0, 1, 3])
def test_reference_test_scenes_known_answers(gpu, tmp_path, degree):
    """tests/test_ply_loader.cpp:55-73: Gaussian 1 at (2, 3, 4), scale (1, 2, 0.5), opacity 0.5,
    rotation (1, 0, 0, 0); SH coefficient 1 of red/green/blue = f_rest 0 / per / 2 per."""
    path = tmp_path / f"ref{degree}.ply"
    path.write_bytes(reference_test_scene(degree, 3))
    with GpuRenderer.from_ply(str(path)) as r:
        a = r.scene_arrays()
        check_equal_to_reference(r, path)
    assert np.allclose(a["pos"][1], [2, 3, 4]) and np.allclose(a["scale"][1], [1, 2, 0.5], atol=1e-6)
    assert abs(a["opacity"][1] - 0.5) < 1e-6 and np.allclose(a["rot"][1], [1, 0, 0, 0])
    per = (degree + 1) ** 2 - 1
    assert np.allclose(a["sh"][1, 0], [0.1, 0.2, 0.3])
    if per:
        assert np.allclose(a["sh"][1, 1], [100, 100 + per, 100 + 2 * per])
        assert a["sh"][1, per, 2] == 100 + 3 * per - 1


@needs_ref
@pytest.mark.parametrize("case", sorted(valid_variants()))
def test_header_variants_match_reference_loader(gpu, tmp_path, case):
    path = tmp_path / f"{case}.ply"
    path.write_bytes(valid_variants()[case])
    with GpuRenderer.from_ply(str(path)) as r:
        check_equal_to_reference(r, path)


@needs_ref
def test_degree_4_file_keeps_degree_3(gpu, tmp_path):
    P = standard_props(4)
    rows = np.random.default_rng(4).normal(0, 1, (300, len(P)))
    path = tmp_path / "deg4.ply"
    path.write_bytes(ply_file(P, rows))
    with GpuRenderer.from_ply(str(path)) as r:
        assert r.sh_degree == 3
        check_equal_to_reference(r, path)


@needs_ref
def test_streamed_load_across_chunks(gpu, tmp_path):
    """300,000 SH3 Gaussians: a 74 MB body read in three 32 MB chunks through the pinned pair."""
    sc = scenes.synthetic_scene(300_000, 1280, 720, seed=11)
    path = tmp_path / "big.ply"
    scenes.write_ply(path, sc)
    print("case streamed load 300k", flush=True)
    with GpuRenderer.from_ply(str(path)) as r:
        check_equal_to_reference(r, path)


def test_ply_renders_equal_host_activated_renders(gpu):
    sc = scenes.synthetic_scene(200_000, 640, 360, seed=12)
    cam = scenes.pinhole_camera(640, 360)
    with GpuRenderer.from_ply(scenes.ply_bytes(sc)) as a, GpuRenderer(sc) as b:
        for opts in (RenderOptions(), RenderOptions(sh_degree=1, tile=(8, 8), splat_scale=0.7)):
            assert np.array_equal(a.render(cam, None, opts), b.render(cam, None, opts))


def test_missing_file_and_malformed_bytes(gpu, tmp_path):
    with pytest.raises(ValueError, match="cannot open"):
        GpuRenderer.from_ply(str(tmp_path / "none.ply"))
    with pytest.raises(ValueError, match="only binary_little_endian"):
        GpuRenderer.from_ply(b"ply\nformat ascii 1.0\nelement vertex 0\nend_header\n")


def normalized(q):
    q = np.asarray(q, np.float32)
    length = np.sqrt(((q[:, 0] * q[:, 0] + q[:, 1] * q[:, 1]) + q[:, 2] * q[:, 2]) + q[:, 3] * q[:, 3])
    return (q / np.maximum(length, np.float32(1e-12))[:, None]).astype(np.float32)


@pytest.mark.parametrize("degree", [0, 3])
def test_writer_layout_and_round_trip(gpu, tmp_path, degree):
    sc = scenes.synthetic_scene(3000, 320, 240, seed=20 + degree)
    sc.sh = np.ascontiguousarray(sc.sh[:, : (degree + 1) ** 2, :])
    sc.quats = (sc.quats * np.float32(3.0)).astype(np.float32)  # unnormalized, as an optimizer holds them
    sc.quats[0] = 0.0
    t = lambda a: torch.tensor(a, dtype=torch.float32, device="cuda")
    data = save_ply(t(sc.means), t(sc.log_scales), t(sc.quats), t(sc.opacity_logits), t(sc.sh))
    expect = scenes.RawScene(sc.means, sc.log_scales, normalized(sc.quats), sc.opacity_logits, sc.sh)
    assert data == scenes.ply_bytes(expect)
    if oracle.ref_available():
        path = tmp_path / "written.ply"
        path.write_bytes(data)
        ref, _ = oracle.ref_load_ply(path)
        assert np.array_equal(bits(ref["pos"]), bits(sc.means))
        assert np.array_equal(bits(ref["sh"]), bits(sc.sh))
