"""ctypes binding of libaniso_amd.so, the C ABI in include/aniso_amd.h.

The library is built in-tree by ``python -m aniso_amd.build`` (or ``__graft_entry__.build()``).
There is no fallback: if the library is missing, loading raises.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libaniso_amd.so")

P = ctypes.c_void_p


class Camera_(ctypes.Structure):
    _fields_ = [("R", ctypes.c_double * 9), ("t", ctypes.c_double * 3), ("fx", ctypes.c_double),
                ("fy", ctypes.c_double), ("cx", ctypes.c_double), ("cy", ctypes.c_double),
                ("width", ctypes.c_int32), ("height", ctypes.c_int32)]


class Options_(ctypes.Structure):
    _fields_ = [("sh_degree", ctypes.c_int32), ("splat_scale", ctypes.c_float), ("tile_w", ctypes.c_int32),
                ("tile_h", ctypes.c_int32), ("batch_depth", ctypes.c_int32), ("flags", ctypes.c_int32)]


FLAG_NO_CULL = 1
FLAG_NO_GRAPH = 2


class Stats_(ctypes.Structure):
    _fields_ = [("visible", ctypes.c_int64), ("tile_pairs", ctypes.c_int64), ("project_ms", ctypes.c_double),
                ("sort_ms", ctypes.c_double), ("blend_ms", ctypes.c_double), ("pair_capacity", ctypes.c_int64),
                ("retries", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class Params_(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("sh_degree", ctypes.c_int32), ("means", ctypes.c_void_p),
                ("log_scales", ctypes.c_void_p), ("quats", ctypes.c_void_p), ("opacity_logits", ctypes.c_void_p),
                ("sh", ctypes.c_void_p), ("activated", ctypes.c_int32)]


class Grads_(ctypes.Structure):
    _fields_ = [("means", ctypes.c_void_p), ("log_scales", ctypes.c_void_p), ("quats", ctypes.c_void_p),
                ("opacity_logits", ctypes.c_void_p), ("sh", ctypes.c_void_p)]


class PathStats_(ctypes.Structure):
    _fields_ = [("frames", ctypes.c_int64), ("chunks", ctypes.c_int64), ("gpu_ms", ctypes.c_double),
                ("total_ms", ctypes.c_double)]


# ans_frame_sink: int (*)(void* user, int64_t frame, const uint8_t* rgb, size_t bytes)
FrameSink = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_uint8),
                             ctypes.c_size_t)


# Every symbol include/aniso_amd.h declares, with (restype, argtypes).
SIGNATURES = {
    "ans_available": (ctypes.c_int, []),
    "ans_last_error": (ctypes.c_char_p, []),
    "ans_renderer_create": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int32, P, P, P, P, P, ctypes.POINTER(P)]),
    "ans_renderer_create_raw": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int32, P, P, P, P, P, ctypes.POINTER(P)]),
    "ans_renderer_destroy": (None, [P]),
    "ans_ply_inspect": (ctypes.c_int, [P, ctypes.c_size_t, ctypes.POINTER(ctypes.c_int64),
                                       ctypes.POINTER(ctypes.c_int32)]),
    "ans_renderer_create_ply": (ctypes.c_int, [P, ctypes.c_size_t, ctypes.POINTER(P)]),
    "ans_renderer_load_ply": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(P)]),
    "ans_renderer_scene": (ctypes.c_int, [P, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int32), P, P,
                                          P, P, P]),
    "ans_render": (ctypes.c_int, [P, ctypes.POINTER(Camera_), ctypes.POINTER(Options_), P, ctypes.c_size_t,
                                  ctypes.POINTER(Stats_)]),
    "ans_render_float": (ctypes.c_int, [P, ctypes.POINTER(Camera_), ctypes.POINTER(Options_), P, P,
                                        ctypes.c_size_t, ctypes.POINTER(Stats_)]),
    "ans_render_async": (ctypes.c_int, [P, ctypes.POINTER(Camera_), ctypes.POINTER(Options_), P, P]),
    "ans_render_wait": (ctypes.c_int, [P, ctypes.POINTER(Stats_)]),
    "ans_frame_counts": (ctypes.c_int, [P, P, P, P, P]),
    "ans_frame_dump": (ctypes.c_int, [P, P, P, P, P, P]),
    "ans_activate": (None, [ctypes.c_int64, P, P, P, P, P, P]),
    "ans_render_views": (ctypes.c_int, [P, ctypes.c_int32, ctypes.POINTER(Camera_), ctypes.POINTER(Options_), P, P,
                                        ctypes.POINTER(Stats_)]),
    "ans_render_views_async": (ctypes.c_int, [P, ctypes.c_int32, ctypes.POINTER(Camera_), ctypes.POINTER(Options_),
                                              P, P]),
    "ans_views_wait": (ctypes.c_int, [P, ctypes.POINTER(Stats_)]),
    "ans_views_autotune": (ctypes.c_int, [P, ctypes.c_int32, ctypes.POINTER(Camera_), ctypes.POINTER(Options_), P, P,
                                          ctypes.c_int32, ctypes.POINTER(Options_), ctypes.POINTER(ctypes.c_double)]),
    "ans_render_path": (ctypes.c_int, [P, ctypes.c_int64, ctypes.POINTER(Camera_), P, ctypes.POINTER(Options_),
                                       ctypes.c_int32, FrameSink, P, ctypes.POINTER(PathStats_)]),
    "ans_views_frame_counts": (ctypes.c_int, [P, ctypes.c_int32, P, P, P]),
    "ans_views_frame_dump": (ctypes.c_int, [P, ctypes.c_int32, P, P]),
    "ans_diff_create": (ctypes.c_int, [ctypes.POINTER(P)]),
    "ans_diff_destroy": (None, [P]),
    "ans_diff_forward": (ctypes.c_int, [P, ctypes.POINTER(Params_), ctypes.POINTER(Camera_), ctypes.POINTER(Options_),
                                        P, P, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(Stats_)]),
    "ans_diff_backward": (ctypes.c_int, [P, ctypes.c_int64, ctypes.POINTER(Params_), P, ctypes.POINTER(Grads_), P]),
    "ans_diff_release": (ctypes.c_int, [P, ctypes.c_int64]),
    "ans_diff_splat_grads": (ctypes.c_int, [P, ctypes.c_int64, P]),
    "ans_diff_radii": (ctypes.c_int, [P, ctypes.c_int64, P, P]),
    "ans_diff_screen_grads": (ctypes.c_int, [P, ctypes.c_int64, P, P]),
    "ans_diff_frame_counts": (ctypes.c_int, [P, ctypes.c_int64, P, P, P]),
    "ans_diff_frame_dump": (ctypes.c_int, [P, ctypes.c_int64, P, P]),
    "ans_diff_views_forward": (ctypes.c_int, [P, ctypes.POINTER(Params_), ctypes.c_int32, ctypes.POINTER(Camera_),
                                              ctypes.POINTER(Options_), P, P, ctypes.POINTER(ctypes.c_int64),
                                              ctypes.POINTER(Stats_)]),
    "ans_diff_views_backward": (ctypes.c_int, [P, ctypes.c_int64, ctypes.POINTER(Params_), P, ctypes.POINTER(Grads_),
                                               P, P]),
    "ans_diff_views_alpha": (ctypes.c_int, [P, ctypes.c_int64, P, P]),
    "ans_diff_views_radii": (ctypes.c_int, [P, ctypes.c_int64, P, P]),
    "ans_diff_views_counts": (ctypes.c_int, [P, ctypes.c_int64, ctypes.c_int32, P, P, P]),
    "ans_diff_views_dump": (ctypes.c_int, [P, ctypes.c_int64, ctypes.c_int32, P, P]),
    "ans_diff_forward_noalloc": (ctypes.c_int, [ctypes.POINTER(Params_), ctypes.POINTER(Camera_),
                                                ctypes.POINTER(Options_), P, P, ctypes.POINTER(Stats_)]),
    "ans_ply_write": (ctypes.c_int, [ctypes.POINTER(Params_), P, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t), P]),
    "ans_render_autotune": (ctypes.c_int, [P, ctypes.POINTER(Camera_), ctypes.POINTER(Options_), P, ctypes.c_int32,
                                           ctypes.POINTER(Options_), ctypes.POINTER(ctypes.c_double)]),
    "ans_colmap_load": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(Camera_), P, ctypes.c_size_t, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_int)]),
    "ans_png_encode": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, P, ctypes.c_size_t,
                                      ctypes.POINTER(ctypes.c_size_t)]),
    "ans_png_write": (ctypes.c_int, [ctypes.c_char_p, P, ctypes.c_int, ctypes.c_int]),
    "ans_png_device_sizes": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_size_t),
                                            ctypes.POINTER(ctypes.c_size_t)]),
    "ans_png_encode_device": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, P, ctypes.c_size_t, P, ctypes.c_size_t,
                                             P]),
}

_lib = None


def load():
    """Loads libaniso_amd.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m aniso_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != 0:
        msg = load().ans_last_error().decode(errors="replace")
        if rc == 1:
            raise ValueError(msg)
        raise RuntimeError(msg)
