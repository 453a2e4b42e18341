"""COLMAP cameras in, PNG frames out: the files the reference CLI's verbs read and write around
the renderer (src/camera.cpp:59-140, src/image.cpp:41-85), through the library's C ABI.

    names, cams = load_colmap("scene/sparse/0")      # one scenes.Camera per image, by name
    write_png("frame.png", rgb)                       # rgb: uint8 [H, W, 3]
    data = png_bytes(rgb)                             # the same file image as bytes
    dev = png_device(frame)                           # torch uint8 CUDA tensor [H, W, 3] -> file image on the GPU
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _lib
from .scenes import Camera

NAME_CAPACITY = 4096


def load_colmap(sparse_dir) -> tuple[list[str], list[Camera]]:
    """aniso::loadColmapCameras (camera.h:44): image names and cameras, sorted by name. Raises
    ValueError with the reference's message for missing or malformed files."""
    L = _lib.load()
    path = os.fsencode(os.fspath(sparse_dir))
    count = ctypes.c_int()
    _lib.check(L.ans_colmap_load(path, None, None, 0, 0, ctypes.byref(count)))
    n = count.value
    cams = (_lib.Camera_ * max(n, 1))()
    names = ctypes.create_string_buffer(NAME_CAPACITY * max(n, 1))
    _lib.check(L.ans_colmap_load(path, cams, names, NAME_CAPACITY, n, ctypes.byref(count)))
    out_names, out = [], []
    for i in range(n):
        raw = names.raw[i * NAME_CAPACITY:(i + 1) * NAME_CAPACITY]
        out_names.append(raw[:raw.index(b"\0")].decode(errors="surrogateescape"))
        c = cams[i]
        out.append(Camera(np.array(c.R[:], np.float64).reshape(3, 3), np.array(c.t[:], np.float64), c.fx, c.fy, c.cx,
                          c.cy, c.width, c.height))
    return out_names, out


def _rgb(rgb) -> np.ndarray:
    a = np.ascontiguousarray(rgb, dtype=np.uint8)
    if a.ndim != 3 or a.shape[2] != 3:
        raise ValueError("rgb must be [height, width, 3] uint8")
    return a


def png_bytes(rgb) -> bytes:
    """The PNG file image aniso::writePng writes for rgb [H, W, 3] uint8."""
    a = _rgb(rgb)
    h, w = a.shape[:2]
    L = _lib.load()
    size = ctypes.c_size_t()
    _lib.check(L.ans_png_encode(a.ctypes.data_as(ctypes.c_void_p), w, h, None, 0, ctypes.byref(size)))
    out = ctypes.create_string_buffer(size.value)
    _lib.check(L.ans_png_encode(a.ctypes.data_as(ctypes.c_void_p), w, h, out, size.value, ctypes.byref(size)))
    return out.raw[:size.value]


def write_png(path, rgb) -> None:
    """aniso::writePng (image.h:22)."""
    a = _rgb(rgb)
    h, w = a.shape[:2]
    _lib.check(_lib.load().ans_png_write(os.fsencode(os.fspath(path)), a.ctypes.data_as(ctypes.c_void_p), w, h))


def png_device(frame, out=None, workspace=None, stream=None):
    """The PNG file image of a CUDA uint8 tensor [H, W, 3], encoded on the GPU into a CUDA uint8
    tensor (out, if given, must be large enough; the returned tensor is its prefix). Enqueued on
    torch's current stream (or `stream`, a raw HIP stream handle) without waiting."""
    import torch

    if frame.dim() != 3 or frame.shape[2] != 3 or frame.dtype != torch.uint8 or not frame.is_cuda:
        raise ValueError("frame must be a CUDA uint8 tensor [height, width, 3]")
    frame = frame.contiguous()
    h, w = int(frame.shape[0]), int(frame.shape[1])
    L = _lib.load()
    fb, wb = ctypes.c_size_t(), ctypes.c_size_t()
    _lib.check(L.ans_png_device_sizes(w, h, ctypes.byref(fb), ctypes.byref(wb)))
    if out is None:
        out = torch.empty(fb.value, dtype=torch.uint8, device=frame.device)
    if workspace is None:
        workspace = torch.empty(wb.value, dtype=torch.uint8, device=frame.device)
    if out.numel() < fb.value or workspace.numel() < wb.value:
        raise ValueError("out or workspace too small")
    s = stream if stream is not None else torch.cuda.current_stream(frame.device).cuda_stream
    _lib.check(L.ans_png_encode_device(ctypes.c_void_p(frame.data_ptr()), w, h, ctypes.c_void_p(out.data_ptr()),
                                       out.numel(), ctypes.c_void_p(workspace.data_ptr()), workspace.numel(),
                                       ctypes.c_void_p(s)))
    return out[:fb.value]
