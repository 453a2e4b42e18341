"""The FP64 gradient oracle (oracle/grad.cpp) is pinned by FP64 central finite differences of
its own forward: an analytic adjoint with a dropped term, a wrong sign or a transposed operand
fails this by orders of magnitude. The forward it differentiates is pinned to the reference's
image by the oracle tests."""
import numpy as np

import oracle
from aniso_amd import scenes


def small_case(seed=3):
    raw = scenes.synthetic_scene(3000, 160, 96, seed=seed)
    raw.log_scales += 1.8  # larger splats: more pairs per Gaussian
    cam = scenes.pose_camera([0.1, 0.0, -1.0], [0.05, 0.0, 1.0], [0, -1, 0], 160, 96, 60.0)
    return raw, cam


def rel_errors(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def test_fp64_adjoint_matches_central_differences():
    raw, cam = small_case()
    rng = np.random.default_rng(0)
    img, _, _ = oracle.grad_run(raw, cam, np.zeros((cam.height, cam.width, 3)))
    dLdC = np.sign(img - (img + rng.normal(0, 0.05, img.shape)))  # L1 against a perturbed target
    _, grad, vis = oracle.grad_run(raw, cam, dLdC)
    assert vis > 1000
    idx = rng.choice(np.flatnonzero(np.abs(grad).sum(1) > 0), 60, replace=False)
    fd, flags = oracle.grad_fd(raw, cam, dLdC, idx)
    ok = flags == 0
    assert ok.mean() > 0.95
    for name, sl in oracle.PARAM_GROUPS.items():
        a, b = grad[idx][:, sl][ok[:, sl]], fd[:, sl][ok[:, sl]]
        err = rel_errors(a, b)
        print(f"{name:15s} samples {a.size:5d}  relative error {err:.2e}")
        assert err < 1e-5, name


def test_fp64_adjoint_is_linear_in_the_image_gradient():
    raw, cam = small_case(5)
    rng = np.random.default_rng(1)
    g1 = rng.normal(size=(cam.height, cam.width, 3))
    g2 = rng.normal(size=(cam.height, cam.width, 3))
    _, a, _ = oracle.grad_run(raw, cam, g1)
    _, b, _ = oracle.grad_run(raw, cam, g2)
    _, c, _ = oracle.grad_run(raw, cam, 2 * g1 - 3 * g2)
    assert np.allclose(c, 2 * a - 3 * b, rtol=1e-9, atol=1e-12 * np.abs(c).max())


def test_fp64_forward_matches_float_oracle_image():
    raw, cam = small_case(7)
    img, _, _ = oracle.grad_run(raw, cam, np.zeros((cam.height, cam.width, 3)))
    act = oracle.activate(raw)
    with oracle.Frame(act, cam, oracle.KEYEXACT, tile=(16, 16)) as f:
        ref = f.image_float()
    assert np.abs(img - ref).max() < 2e-3


def test_culled_gaussians_get_zero_gradient():
    raw, cam = small_case(9)
    raw.means[:50, 2] = -5.0  # behind the camera
    _, grad, _ = oracle.grad_run(raw, cam, np.ones((cam.height, cam.width, 3)))
    assert not grad[:50].any()
