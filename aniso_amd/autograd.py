"""Differentiable rasterization for PyTorch: a torch.autograd.Function over the C ABI's
ans_diff_* entry points (hand-written HIP forward and backward kernels).

    image = rasterize(means, log_scales, quats, opacity_logits, sh, camera, options)
    loss = (image - target).abs().mean(); loss.backward()

Parameters are the raw values an optimizer holds (the 3DGS PLY layout, pipeline/train_gsplat.py:85-92):
means [N,3], log_scales [N,3], quats [N,4] (w, x, y, z; normalized inside), opacity_logits [N],
sh [N,(d+1)^2,3]. The image is the colour before the [0, 1] clamp, float32 [H, W, 3].
This stands where the reference's training script calls gsplat's rasterization
(pipeline/train_gsplat.py:96-102) with exp/sigmoid/normalize applied by the caller; here the
activations are part of the differentiated function.
"""
from __future__ import annotations

import ctypes
import threading

import numpy as np
import torch

from . import _lib
from .renderer import RenderOptions, RenderStats, camera_c
from .scenes import Camera

_handles = {}
_lock = threading.Lock()


def _handle(device_index: int):
    """One ans_diff handle per device (created on first use, with that device current)."""
    with _lock:
        h = _handles.get(device_index)
        if h is None:
            with torch.cuda.device(device_index):
                hp = ctypes.c_void_p()
                _lib.check(_lib.load().ans_diff_create(ctypes.byref(hp)))
            h = _handles[device_index] = hp
        return h


def _params(means, log_scales, quats, opacity_logits, sh, activated: bool = False) -> _lib.Params_:
    n = means.shape[0]
    degree = int(round(sh.shape[1] ** 0.5)) - 1
    return _lib.Params_(n, degree, means.data_ptr(), log_scales.data_ptr(), quats.data_ptr(),
                        opacity_logits.data_ptr(), sh.data_ptr(), 1 if activated else 0)


def _check(t, name, shape):
    if not t.is_cuda or t.dtype != torch.float32:
        raise ValueError(f"{name} must be a float32 CUDA tensor")
    if tuple(t.shape[1:]) != tuple(shape[1:]) or t.shape[0] != shape[0]:
        raise ValueError(f"{name} has shape {tuple(t.shape)}, expected {shape}")


class _Frame:
    """A forward frame kept on the device for its backward pass; released when dropped."""

    def __init__(self, handle, fid):
        self.handle, self.fid = handle, fid

    def release(self):
        if self.fid is not None:
            _lib.load().ans_diff_release(self.handle, self.fid)
            self.fid = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class RasterizeFunction(torch.autograd.Function):
    """Inputs: the five raw parameter tensors, an optional `means2d` leaf ([N, 2], only its
    gradient is used: it receives dL/d(u, v), the screen-space mean gradient), the camera, the
    options, optional stats. Outputs: the image and the per-Gaussian radii (not differentiable;
    empty unless means2d is given)."""

    @staticmethod
    def forward(ctx, means, log_scales, quats, opacity_logits, sh, means2d, camera: Camera, options: RenderOptions,
                stats: RenderStats | None):
        n = means.shape[0]
        k = sh.shape[1]
        for t, name, shape in ((means, "means", (n, 3)), (log_scales, "log_scales", (n, 3)), (quats, "quats", (n, 4)),
                               (opacity_logits, "opacity_logits", (n,)), (sh, "sh", (n, k, 3))):
            _check(t, name, shape)
        if means2d is not None:
            _check(means2d, "means2d", (n, 2))
        means, log_scales, quats, opacity_logits, sh = (t.contiguous() for t in
                                                        (means, log_scales, quats, opacity_logits, sh))
        dev = means.device.index if means.device.index is not None else torch.cuda.current_device()
        handle = _handle(dev)
        L = _lib.load()
        image = torch.empty((camera.height, camera.width, 3), dtype=torch.float32, device=means.device)
        fid = ctypes.c_int64()
        st = _lib.Stats_() if stats is not None else None  # stats wait for the frame to finish
        _lib.check(L.ans_diff_forward(handle, ctypes.byref(_params(means, log_scales, quats, opacity_logits, sh)),
                                      ctypes.byref(camera_c(camera)), ctypes.byref(options.c()), image.data_ptr(),
                                      _stream(), ctypes.byref(fid), ctypes.byref(st) if st is not None else None))
        frame = _Frame(handle, fid.value)
        # Radii only for the info path (means2d given); a plain render skips their launch.
        radii = torch.empty(n if means2d is not None else 0, dtype=torch.int32, device=means.device)
        if means2d is not None:
            _lib.check(L.ans_diff_radii(handle, fid.value, radii.data_ptr(), _stream()))
        if stats is not None:
            for f in ("visible", "tile_pairs", "project_ms", "sort_ms", "blend_ms", "pair_capacity", "retries"):
                setattr(stats, f, getattr(st, f))
        # Inside forward, grad mode is off; needs_input_grad says whether a backward will come.
        if any(ctx.needs_input_grad[:6]):
            ctx.frame = frame
            ctx.save_for_backward(means, log_scales, quats, opacity_logits, sh)
        else:
            frame.release()
        ctx.mark_non_differentiable(radii)
        return image, radii

    @staticmethod
    def backward(ctx, grad_image, grad_radii):
        means, log_scales, quats, opacity_logits, sh = ctx.saved_tensors
        grads = [torch.empty_like(t) for t in (means, log_scales, quats, opacity_logits, sh)]
        g = grad_image.contiguous().float()
        frame = ctx.frame
        L = _lib.load()
        _lib.check(L.ans_diff_backward(frame.handle, frame.fid,
                                       ctypes.byref(_params(means, log_scales, quats, opacity_logits, sh)),
                                       g.data_ptr(), ctypes.byref(_lib.Grads_(*[t.data_ptr() for t in grads])),
                                       _stream()))
        grad_uv = None
        if ctx.needs_input_grad[5]:
            grad_uv = torch.empty((means.shape[0], 2), dtype=torch.float32, device=means.device)
            _lib.check(L.ans_diff_screen_grads(frame.handle, frame.fid, grad_uv.data_ptr(), _stream()))
        # Releasing on the host is safe: the device buffers go back to a pool that is only reused
        # by later work on the same stream.
        frame.release()
        return (*grads, grad_uv, None, None, None)


def rasterize(means, log_scales, quats, opacity_logits, sh, camera: Camera, options: RenderOptions | None = None,
              stats: RenderStats | None = None, info: bool = False):
    """Differentiable render: unclamped colour image [H, W, 3] (float32).

    With info=True returns (image, info) where info holds what a densification strategy reads
    (gsplat's `info`, pipeline/train_gsplat.py:177, :185): "radii" (int32 [N], pixels, 0 if
    culled), "means2d" (a zero [N, 2] leaf whose .grad receives dL/d(u, v) in pixels after
    backward), "width", "height", "n_cameras" (1)."""
    opts = options or RenderOptions()
    if not info:
        image, _ = RasterizeFunction.apply(means, log_scales, quats, opacity_logits, sh, None, camera, opts, stats)
        return image
    means2d = torch.zeros((means.shape[0], 2), dtype=torch.float32, device=means.device, requires_grad=True)
    image, radii = RasterizeFunction.apply(means, log_scales, quats, opacity_logits, sh, means2d, camera, opts, stats)
    return image, {"radii": radii, "means2d": means2d, "width": camera.width, "height": camera.height,
                   "n_cameras": 1}


def frame_lists(means, log_scales, quats, opacity_logits, sh, camera: Camera, options: RenderOptions | None = None):
    """Per-tile lists (start [T+1], concatenated lists [P]) of a forward frame, for evaluating a
    reference gradient on exactly the GPU's lists."""
    dev = means.device.index if means.device.index is not None else torch.cuda.current_device()
    handle = _handle(dev)
    L = _lib.load()
    image = torch.empty((camera.height, camera.width, 3), dtype=torch.float32, device=means.device)
    fid = ctypes.c_int64()
    _lib.check(L.ans_diff_forward(handle, ctypes.byref(_params(means, log_scales, quats, opacity_logits, sh)),
                                  ctypes.byref(camera_c(camera)), ctypes.byref((options or RenderOptions()).c()),
                                  image.data_ptr(), _stream(), ctypes.byref(fid), None))
    try:
        vis, pairs, tiles = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _lib.check(L.ans_diff_frame_counts(handle, fid.value, ctypes.byref(vis), ctypes.byref(pairs),
                                           ctypes.byref(tiles)))
        values = np.empty(pairs.value, np.uint32)
        ranges = np.empty((tiles.value, 2), np.uint32)
        _lib.check(L.ans_diff_frame_dump(handle, fid.value, values.ctypes.data_as(ctypes.c_void_p),
                                         ranges.ctypes.data_as(ctypes.c_void_p)))
    finally:
        L.ans_diff_release(handle, fid.value)
    # Ranges of empty tiles are [0, 0); rebuild starts from the lengths in tile order.
    lengths = (ranges[:, 1] - ranges[:, 0]).astype(np.int64)
    start = np.zeros(len(ranges) + 1, np.uint32)
    start[1:] = np.cumsum(lengths)
    return start, values, image


def forward_backward(means, log_scales, quats, opacity_logits, sh, camera: Camera, grad_image,
                     options: RenderOptions | None = None):
    """One forward and backward without autograd: returns (image, [grads of the five parameter
    tensors], splat gradients [n, 9] = dL/d{u, v, conic A, B, C, opacity, r, g, b} as a host
    array; u, v are the screen-space mean gradients densification heuristics use)."""
    dev = means.device.index if means.device.index is not None else torch.cuda.current_device()
    handle = _handle(dev)
    L = _lib.load()
    params = _params(means, log_scales, quats, opacity_logits, sh)
    image = torch.empty((camera.height, camera.width, 3), dtype=torch.float32, device=means.device)
    fid = ctypes.c_int64()
    _lib.check(L.ans_diff_forward(handle, ctypes.byref(params), ctypes.byref(camera_c(camera)),
                                  ctypes.byref((options or RenderOptions()).c()), image.data_ptr(), _stream(),
                                  ctypes.byref(fid), None))
    try:
        grads = [torch.empty_like(t) for t in (means, log_scales, quats, opacity_logits, sh)]
        g = grad_image.contiguous().float()
        _lib.check(L.ans_diff_backward(handle, fid.value, ctypes.byref(params), g.data_ptr(),
                                       ctypes.byref(_lib.Grads_(*[t.data_ptr() for t in grads])), _stream()))
        splat = np.empty((means.shape[0], 9), np.float32)
        _lib.check(L.ans_diff_splat_grads(handle, fid.value, splat.ctypes.data_as(ctypes.c_void_p)))
    finally:
        L.ans_diff_release(handle, fid.value)
    return image, grads, splat


def save_ply(means, log_scales, quats, opacity_logits, sh, path=None) -> bytes:
    """The raw parameters (float32 CUDA tensors, as for `rasterize`) as a 3DGS PLY file image,
    the layout of the training pipeline's writer (pipeline/train_gsplat.py:109-128); the rows are
    interleaved on the device (ans_ply_write) on torch's current stream. Writes `path` if given
    and returns the bytes."""
    n = means.shape[0]
    K = sh.shape[1]
    for t, name, shape in ((means, "means", (n, 3)), (log_scales, "log_scales", (n, 3)), (quats, "quats", (n, 4)),
                           (opacity_logits, "opacity_logits", (n,)), (sh, "sh", (n, K, 3))):
        _check(t, name, shape)
    tensors = [t.contiguous() for t in (means, log_scales, quats, opacity_logits, sh)]
    p = _params(*tensors)
    size = ctypes.c_size_t()
    L = _lib.load()
    _lib.check(L.ans_ply_write(ctypes.byref(p), None, 0, ctypes.byref(size), None))
    out = bytearray(size.value)
    buf = (ctypes.c_char * len(out)).from_buffer(out)
    with torch.cuda.device(means.device):
        _lib.check(L.ans_ply_write(ctypes.byref(p), buf, len(out), ctypes.byref(size), _stream()))
    data = bytes(out)
    if path is not None:
        with open(path, "wb") as f:
            f.write(data)
    return data


# ---- Batched views: many cameras, one parameter set ------------------------------------------

def _device_index(t):
    return t.device.index if t.device.index is not None else torch.cuda.current_device()


class RasterizeViewsFunction(torch.autograd.Function):
    """Inputs: the five parameter tensors (raw, or with activated=True standard deviations and
    opacities in place of log-scales and logits), an optional `means2d` leaf ([V, N, 2], receives
    dL/d(u, v) per view), the cameras, options, activated, optional stats list, whether to form the
    alpha images. Outputs: all views' images concatenated (flat float32, view v's [H_v, W_v, 3] at
    its offset), the radii [V, N] and the alpha images concatenated (flat, 1 - final transmittance;
    empty unless asked for; neither differentiable)."""

    @staticmethod
    def forward(ctx, means, log_scales, quats, opacity_logits, sh, means2d, cameras, options: RenderOptions,
                activated: bool, stats, want_alpha: bool = True):
        n = means.shape[0]
        k = sh.shape[1]
        V = len(cameras)
        for t, name, shape in ((means, "means", (n, 3)), (log_scales, "log_scales", (n, 3)), (quats, "quats", (n, 4)),
                               (opacity_logits, "opacity_logits", (n,)), (sh, "sh", (n, k, 3))):
            _check(t, name, shape)
        if means2d is not None:
            _check(means2d, "means2d", (V, n, 2))
        means, log_scales, quats, opacity_logits, sh = (t.contiguous() for t in
                                                        (means, log_scales, quats, opacity_logits, sh))
        handle = _handle(_device_index(means))
        L = _lib.load()
        pixels = [c.width * c.height for c in cameras]
        offs = np.concatenate([[0], np.cumsum(pixels)]).astype(np.int64)
        flat = torch.empty(int(offs[-1]) * 3, dtype=torch.float32, device=means.device)
        alpha = torch.empty(int(offs[-1]) if want_alpha else 0, dtype=torch.float32, device=means.device)
        cams = (_lib.Camera_ * V)(*[camera_c(c) for c in cameras])
        imgs = (ctypes.c_void_p * V)(*[flat.data_ptr() + 12 * int(o) for o in offs[:-1]])
        st = (_lib.Stats_ * V)()
        fid = ctypes.c_int64()
        params = _params(means, log_scales, quats, opacity_logits, sh, activated)
        _lib.check(L.ans_diff_views_forward(handle, ctypes.byref(params), V, cams, ctypes.byref(options.c()), imgs,
                                            _stream(), ctypes.byref(fid), st))
        frame = _Frame(handle, fid.value)
        if want_alpha:
            alps = (ctypes.c_void_p * V)(*[alpha.data_ptr() + 4 * int(o) for o in offs[:-1]])
            _lib.check(L.ans_diff_views_alpha(handle, fid.value, alps, _stream()))
        radii = torch.empty((V, n), dtype=torch.int32, device=means.device)
        if means2d is not None:
            _lib.check(L.ans_diff_views_radii(handle, fid.value, radii.data_ptr(), _stream()))
        if stats is not None:
            stats.clear()
            for s_ in st:
                stats.append(RenderStats(s_.visible, s_.tile_pairs, s_.project_ms, s_.sort_ms, s_.blend_ms,
                                         s_.pair_capacity, s_.retries))
        if any(ctx.needs_input_grad[:6]):
            ctx.frame = frame
            ctx.offs = offs
            ctx.activated = activated
            ctx.save_for_backward(means, log_scales, quats, opacity_logits, sh)
        else:
            frame.release()
        ctx.mark_non_differentiable(radii, alpha)
        return flat, radii, alpha

    @staticmethod
    def backward(ctx, grad_flat, grad_radii, grad_alpha):
        means, log_scales, quats, opacity_logits, sh = ctx.saved_tensors
        grads = [torch.empty_like(t) for t in (means, log_scales, quats, opacity_logits, sh)]
        g = grad_flat.contiguous().float()
        V = len(ctx.offs) - 1
        gptrs = (ctypes.c_void_p * V)(*[g.data_ptr() + 12 * int(o) for o in ctx.offs[:-1]])
        grad_uv = None
        if ctx.needs_input_grad[5]:
            grad_uv = torch.empty((V, means.shape[0], 2), dtype=torch.float32, device=means.device)
        frame = ctx.frame
        L = _lib.load()
        _lib.check(L.ans_diff_views_backward(
            frame.handle, frame.fid, ctypes.byref(_params(means, log_scales, quats, opacity_logits, sh, ctx.activated)),
            gptrs, ctypes.byref(_lib.Grads_(*[t.data_ptr() for t in grads])),
            grad_uv.data_ptr() if grad_uv is not None else None, _stream()))
        frame.release()
        return (*grads, grad_uv, None, None, None, None, None)


def _split_views(flat, cameras, channels):
    out, o = [], 0
    for c in cameras:
        k = c.width * c.height * channels
        out.append(flat[o:o + k].view(c.height, c.width, channels) if channels > 1 else
                   flat[o:o + k].view(c.height, c.width))
        o += k
    return out


def rasterize_views(means, log_scales, quats, opacity_logits, sh, cameras, options: RenderOptions | None = None,
                    stats: list | None = None, info: bool = False, activated: bool = False, flat: bool = False):
    """Differentiable render of many cameras (uneven sizes allowed) in one batched pass: a list of
    unclamped colour images [H_v, W_v, 3] (with flat=True, one tensor holding them all, view after
    view, so that a loss over all views is a few whole-tensor operations). Gradients w.r.t. the
    parameters are summed over the views. With info=True returns (images, info) with "radii"
    (int32 [V, N]), "means2d" (a zero [V, N, 2] leaf whose .grad receives dL/d(u, v) per view),
    "alphas" (list of [H_v, W_v]), "n_cameras"."""
    opts = options or RenderOptions()
    means2d = None
    if info:
        means2d = torch.zeros((len(cameras), means.shape[0], 2), dtype=torch.float32, device=means.device,
                              requires_grad=True)
    flat_images, radii, alpha = RasterizeViewsFunction.apply(means, log_scales, quats, opacity_logits, sh, means2d,
                                                             list(cameras), opts, activated, stats, info)
    images = flat_images if flat else _split_views(flat_images, cameras, 3)
    if not info:
        return images
    return images, {"radii": radii, "means2d": means2d, "alphas": _split_views(alpha, cameras, 1),
                    "n_cameras": len(cameras)}


def views_frame_lists(means, log_scales, quats, opacity_logits, sh, cameras, options: RenderOptions | None = None,
                      activated: bool = False):
    """Per-view (start [T+1], concatenated lists [P]) of a batched differentiable forward, for
    evaluating reference gradients on exactly the GPU's lists."""
    handle = _handle(_device_index(means))
    L = _lib.load()
    V = len(cameras)
    bufs = [torch.empty((c.height, c.width, 3), dtype=torch.float32, device=means.device) for c in cameras]
    cams = (_lib.Camera_ * V)(*[camera_c(c) for c in cameras])
    imgs = (ctypes.c_void_p * V)(*[b.data_ptr() for b in bufs])
    fid = ctypes.c_int64()
    _lib.check(L.ans_diff_views_forward(handle, ctypes.byref(_params(means, log_scales, quats, opacity_logits, sh,
                                                                      activated)),
                                        V, cams, ctypes.byref((options or RenderOptions()).c()), imgs, _stream(),
                                        ctypes.byref(fid), None))
    out = []
    try:
        for v in range(V):
            vis, pairs, tiles = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
            _lib.check(L.ans_diff_views_counts(handle, fid.value, v, ctypes.byref(vis), ctypes.byref(pairs),
                                               ctypes.byref(tiles)))
            values = np.empty(pairs.value, np.uint32)
            ranges = np.empty((tiles.value, 2), np.uint32)
            _lib.check(L.ans_diff_views_dump(handle, fid.value, v, values.ctypes.data_as(ctypes.c_void_p),
                                             ranges.ctypes.data_as(ctypes.c_void_p)))
            lengths = (ranges[:, 1] - ranges[:, 0]).astype(np.int64)
            start = np.zeros(len(ranges) + 1, np.uint32)
            start[1:] = np.cumsum(lengths)
            out.append((start, values))
    finally:
        L.ans_diff_release(handle, fid.value)
    return out, bufs
