"""CPU oracle for the rasterizer -- TEST INFRASTRUCTURE ONLY.

Only the tests, the smoke check and the benchmark's CPU-baseline leg import this package, as the
checker. It wraps two native libraries:

* ``_build/liboracle.so``: this project's C++ restatement of the reference CPU renderer
  (src/render.cpp, src/sh.cpp, src/camera.cpp:12-32, src/ply_loader.cpp:124-139), built from
  ``oracle.cpp`` / ``grad.cpp`` by ``make -C oracle``;
* ``_ref/libaniso_ref.so``: the reference CPU renderer compiled from its own sources by
  ``make -C oracle ref`` (only where the reference tree is present; the built file travels).

Modes: REFERENCE (the reference's arithmetic; reproduces the compiled reference bit for bit at
8x8 tiles), KEYEXACT (FP32 camera transform in the kernel's op order; the bit-exact contract for
the GPU), F64 (double throughout).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libaniso_ref.so")

REFERENCE, KEYEXACT, F64 = 0, 1, 2


class _Camera(ctypes.Structure):
    _fields_ = [("R", ctypes.c_double * 9), ("t", ctypes.c_double * 3), ("fx", ctypes.c_double),
                ("fy", ctypes.c_double), ("cx", ctypes.c_double), ("cy", ctypes.c_double),
                ("width", ctypes.c_int), ("height", ctypes.c_int)]


class _Stats(ctypes.Structure):
    _fields_ = [("visible", ctypes.c_int64), ("pairs", ctypes.c_int64), ("project_ms", ctypes.c_double),
                ("sort_ms", ctypes.c_double), ("blend_ms", ctypes.c_double)]


def build(reference: bool = True) -> None:
    """Compiles the oracle (and the compiled reference when its sources exist)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if reference and os.path.isdir("/root/reference/src"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build(reference=False)
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        L.orc_frame_run.restype = P
        L.orc_frame_run.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_int, P, P, P, P, P,
                                    ctypes.POINTER(_Camera), ctypes.c_int, ctypes.c_float, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int]
        L.orc_frame_free.argtypes = [P]
        L.orc_frame_stats.argtypes = [P, ctypes.POINTER(_Stats)]
        L.orc_frame_tile_count.restype = ctypes.c_int64
        L.orc_frame_tile_count.argtypes = [P]
        L.orc_frame_image.argtypes = [P, P, P]
        L.orc_frame_ncontrib.argtypes = [P, P]
        L.orc_frame_pairs.argtypes = [P, P, P, P]
        L.orc_frame_splats.argtypes = [P, P, P, P, P]
        L.orc_frame_splat_bits.argtypes = [P, P]
        L.orc_activate.argtypes = [ctypes.c_int64, P, P, P, P, P, P]
        L.orc_mt19937_scene.argtypes = [ctypes.c_uint32, ctypes.c_int64, P, P, P, P, P]
        L.orc_grad_run.restype = ctypes.c_int64
        L.orc_grad_run.argtypes = [ctypes.c_int64, P, P, P, P, P, ctypes.POINTER(_Camera), ctypes.c_int,
                                   ctypes.c_double, ctypes.c_int, ctypes.c_int, P, P, P, P, P, P, ctypes.c_int]
        L.orc_grad_fd.argtypes = [ctypes.c_int64, P, P, P, P, P, ctypes.POINTER(_Camera), ctypes.c_int,
                                  ctypes.c_double, ctypes.c_int, ctypes.c_int, P, P, ctypes.c_int64, ctypes.c_double,
                                  P, P, ctypes.c_int]
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref_lib():
    global _ref
    if _ref is None:
        R = ctypes.CDLL(REF_PATH)
        P = ctypes.c_void_p
        R.ref_render.restype = ctypes.c_int
        R.ref_render.argtypes = [ctypes.c_int64, ctypes.c_int, P, P, P, P, P, ctypes.POINTER(_Camera),
                                 ctypes.c_int, ctypes.c_float, P, ctypes.POINTER(_Stats)]
        R.ref_load_ply.restype = P
        R.ref_load_ply.argtypes = [ctypes.c_char_p, P, ctypes.c_size_t, ctypes.POINTER(ctypes.c_double)]
        R.ref_scene_arrays.argtypes = [P, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int), P, P, P, P, P]
        R.ref_scene_free.argtypes = [P]
        R.ref_write_png.restype = ctypes.c_int
        R.ref_write_png.argtypes = [ctypes.c_char_p, P, ctypes.c_int, ctypes.c_int, P, ctypes.c_size_t,
                                    ctypes.POINTER(ctypes.c_double)]
        R.ref_load_colmap.restype = ctypes.c_int
        R.ref_load_colmap.argtypes = [ctypes.c_char_p, ctypes.POINTER(_Camera), P, ctypes.c_size_t, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_int), P, ctypes.c_size_t]
        _ref = R
    return _ref


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _cam(camera) -> _Camera:
    c = _Camera()
    for i, v in enumerate(np.asarray(camera.R, dtype=np.float64).reshape(9)):
        c.R[i] = float(v)
    for i, v in enumerate(np.asarray(camera.t, dtype=np.float64).reshape(3)):
        c.t[i] = float(v)
    c.fx, c.fy, c.cx, c.cy = float(camera.fx), float(camera.fy), float(camera.cx), float(camera.cy)
    c.width, c.height = int(camera.width), int(camera.height)
    return c


class ActivatedScene:
    """Activated float32 arrays as the renderers take them (gaussian.h:19-40)."""

    def __init__(self, pos, scale, rot, opacity, sh):
        f = lambda a: np.ascontiguousarray(a, dtype=np.float32)
        self.pos, self.scale, self.rot, self.opacity, self.sh = f(pos), f(scale), f(rot), f(opacity), f(sh)
        self.n = self.pos.shape[0]
        self.sh_degree = int(round(np.sqrt(self.sh.shape[1]))) - 1


def mt19937_scene(n: int = 20000, seed: int = 7) -> ActivatedScene:
    """The random SH3 scene of tests/test_gpu_render.cpp:41-58 (std::mt19937(7))."""
    pos = np.empty((n, 3), np.float32)
    scale = np.empty((n, 3), np.float32)
    rot = np.empty((n, 4), np.float32)
    opacity = np.empty(n, np.float32)
    sh = np.empty((n, 16, 3), np.float32)
    lib().orc_mt19937_scene(seed, n, _ptr(pos), _ptr(scale), _ptr(rot), _ptr(opacity), _ptr(sh))
    return ActivatedScene(pos, scale, rot, opacity, sh)


def activate(raw) -> ActivatedScene:
    """The reference loader's activations (ply_loader.cpp:124-139), in float with std::exp."""
    n = raw.n
    scale = np.empty((n, 3), np.float32)
    opacity = np.empty(n, np.float32)
    rot = np.empty((n, 4), np.float32)
    ls = np.ascontiguousarray(raw.log_scales, np.float32)
    lg = np.ascontiguousarray(raw.opacity_logits, np.float32)
    q = np.ascontiguousarray(raw.quats, np.float32)
    lib().orc_activate(n, _ptr(ls), _ptr(lg), _ptr(q), _ptr(scale), _ptr(opacity), _ptr(rot))
    return ActivatedScene(raw.means, scale, rot, opacity, raw.sh)


class Frame:
    """One oracle frame; every intermediate is available until close()."""

    def __init__(self, scene: ActivatedScene, camera, mode=KEYEXACT, sh_degree=3, splat_scale=1.0, tile=(16, 16),
                 threads=0):
        self._scene = scene  # keep the arrays alive
        self.mode = mode
        self.camera = camera
        self.tile = tile
        self._h = lib().orc_frame_run(mode, scene.n, scene.sh_degree, _ptr(scene.pos), _ptr(scene.scale),
                                      _ptr(scene.rot), _ptr(scene.opacity), _ptr(scene.sh),
                                      ctypes.byref(_cam(camera)), sh_degree, splat_scale, tile[0], tile[1], threads)
        if not self._h:
            raise ValueError("oracle rejected the frame arguments")
        st = _Stats()
        lib().orc_frame_stats(self._h, ctypes.byref(st))
        self.visible, self.pairs = st.visible, st.pairs
        self.project_ms, self.sort_ms, self.blend_ms = st.project_ms, st.sort_ms, st.blend_ms
        self.tiles = lib().orc_frame_tile_count(self._h)

    def image(self) -> np.ndarray:
        out = np.empty((self.camera.height, self.camera.width, 3), np.uint8)
        lib().orc_frame_image(self._h, _ptr(out), None)
        return out

    def image_float(self) -> np.ndarray:
        out = np.empty((self.camera.height, self.camera.width, 3), np.float64)
        lib().orc_frame_image(self._h, None, _ptr(out))
        return out

    def n_contrib(self) -> np.ndarray:
        out = np.empty((self.camera.height, self.camera.width), np.uint32)
        lib().orc_frame_ncontrib(self._h, _ptr(out))
        return out

    def pairs_arrays(self):
        """(keys u64[P], values u32[P], ranges u32[T,2]) as the GPU sort produces them."""
        keys = np.empty(self.pairs, np.uint64)
        values = np.empty(self.pairs, np.uint32)
        ranges = np.empty((self.tiles, 2), np.uint32)
        lib().orc_frame_pairs(self._h, _ptr(keys), _ptr(values), _ptr(ranges))
        return keys, values, ranges

    def splats(self):
        """(records f64[n,10] = u v A B C opacity depth r g b, rect i32[n,4], radius i32[n], alive u8[n])."""
        n = self._scene.n
        rec = np.empty((n, 10), np.float64)
        rect = np.empty((n, 4), np.int32)
        radius = np.empty(n, np.int32)
        alive = np.empty(n, np.uint8)
        lib().orc_frame_splats(self._h, _ptr(rec), _ptr(rect), _ptr(radius), _ptr(alive))
        return rec, rect, radius, alive

    def splat_bits(self) -> np.ndarray:
        out = np.empty((self._scene.n, 10), np.uint32)
        lib().orc_frame_splat_bits(self._h, _ptr(out))
        return out

    def close(self):
        if self._h:
            lib().orc_frame_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def render(scene: ActivatedScene, camera, mode=REFERENCE, sh_degree=3, splat_scale=1.0, tile=(8, 8), threads=0):
    """RGB8 image [H, W, 3] from the oracle."""
    with Frame(scene, camera, mode, sh_degree, splat_scale, tile, threads) as f:
        return f.image()


def ref_render(scene: ActivatedScene, camera, sh_degree=3, splat_scale=1.0):
    """RGB8 image and stats from the compiled reference CPU renderer (aniso::render)."""
    out = np.empty((camera.height, camera.width, 3), np.uint8)
    st = _Stats()
    rc = ref_lib().ref_render(scene.n, scene.sh_degree, _ptr(scene.pos), _ptr(scene.scale), _ptr(scene.rot),
                              _ptr(scene.opacity), _ptr(scene.sh), ctypes.byref(_cam(camera)), sh_degree,
                              splat_scale, _ptr(out), ctypes.byref(st))
    if rc != 0:
        raise RuntimeError("reference render failed")
    return out, st


class RefPlyError(ValueError):
    """aniso::PlyError raised by the compiled reference loader; the message is its what()."""


def ref_load_ply(path):
    """Loads a PLY with the compiled reference loader (aniso::loadPly). Returns (arrays, ms):
    arrays = dict(pos, scale, rot (w, x, y, z), opacity, sh [n, coeffs, 3], sh_degree) and the
    load's wall time in ms. Raises RefPlyError with the loader's message."""
    R = ref_lib()
    err = ctypes.create_string_buffer(1024)
    ms = ctypes.c_double()
    h = R.ref_load_ply(os.fsencode(path), err, len(err), ctypes.byref(ms))
    if not h:
        raise RefPlyError(err.value.decode(errors="replace"))
    try:
        n, degree = ctypes.c_int64(), ctypes.c_int()
        R.ref_scene_arrays(h, ctypes.byref(n), ctypes.byref(degree), None, None, None, None, None)
        n, degree = n.value, degree.value
        k = (degree + 1) ** 2
        out = dict(pos=np.empty((n, 3), np.float32), scale=np.empty((n, 3), np.float32),
                   rot=np.empty((n, 4), np.float32), opacity=np.empty(n, np.float32),
                   sh=np.empty((n, k, 3), np.float32))
        R.ref_scene_arrays(h, None, None, *[_ptr(out[key]) for key in ("pos", "scale", "rot", "opacity", "sh")])
        out["sh_degree"] = degree
        return out, ms.value
    finally:
        R.ref_scene_free(h)


def ref_write_png(path, rgb: np.ndarray) -> float:
    """aniso::writePng of rgb [H, W, 3] uint8 to path with the compiled reference writer; returns
    its wall time in ms. Raises RuntimeError with the writer's message."""
    rgb = np.ascontiguousarray(rgb, dtype=np.uint8)
    err = ctypes.create_string_buffer(1024)
    ms = ctypes.c_double()
    h, w = rgb.shape[:2]
    if ref_lib().ref_write_png(os.fsencode(path), _ptr(rgb), w, h, err, len(err), ctypes.byref(ms)) != 0:
        raise RuntimeError(err.value.decode(errors="replace"))
    return ms.value


def ref_load_colmap(sparse_dir):
    """aniso::loadColmapCameras with the compiled reference reader: (names, cameras) with cameras
    as dicts of R [3, 3], t [3], fx, fy, cx, cy, width, height. Raises RefPlyError with the
    reader's message (the reference reuses PlyError for bad camera files)."""
    R = ref_lib()
    err = ctypes.create_string_buffer(1024)
    count = ctypes.c_int()
    if R.ref_load_colmap(os.fsencode(sparse_dir), None, None, 0, 0, ctypes.byref(count), err, len(err)) != 0:
        raise RefPlyError(err.value.decode(errors="replace"))
    n = count.value
    cams = (_Camera * max(n, 1))()
    cap = 4096
    names = ctypes.create_string_buffer(cap * max(n, 1))
    if R.ref_load_colmap(os.fsencode(sparse_dir), cams, names, cap, n, ctypes.byref(count), err, len(err)) != 0:
        raise RefPlyError(err.value.decode(errors="replace"))
    out_names, out = [], []
    for i in range(n):
        raw = names.raw[i * cap:(i + 1) * cap]
        out_names.append(raw[:raw.index(b"\0")].decode())
        c = cams[i]
        out.append(dict(R=np.array(c.R[:], np.float64).reshape(3, 3), t=np.array(c.t[:], np.float64), fx=c.fx,
                        fy=c.fy, cx=c.cx, cy=c.cy, width=c.width, height=c.height))
    return out_names, out


# --- FP64 gradients (grad.cpp) ------------------------------------------------------------------

PARAM_GROUPS = {"means": slice(0, 3), "log_scales": slice(3, 6), "quats": slice(6, 10), "opacity_logits": slice(10, 11),
                "sh_band0": slice(11, 14), "sh_band1": slice(14, 23), "sh_band2": slice(23, 38),
                "sh_band3": slice(38, 59)}


def _raw64(raw):
    f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    sh = raw.sh
    if sh.shape[1] != 16:
        full = np.zeros((raw.n, 16, 3), np.float64)
        full[:, :sh.shape[1]] = sh
        sh = full
    return f(raw.means), f(raw.log_scales), f(raw.quats), f(raw.opacity_logits), f(sh)


def grad_run(raw, camera, dLdC, sh_degree=3, splat_scale=1.0, tile=(16, 16), lists=None, threads=0,
             splat_grad=False):
    """FP64 forward image [H,W,3] and analytic gradient [n,59] of L = sum(dLdC * C) w.r.t. the raw
    parameters (means, log-scales, quats, opacity logits, SH[16][3]). `lists` = (start[T+1],
    concatenated per-tile lists) replaces the oracle's own binning."""
    arrays = _raw64(raw)
    n = raw.n
    image = np.empty((camera.height, camera.width, 3), np.float64)
    grad = np.empty((n, 59), np.float64)
    g = np.ascontiguousarray(dLdC, dtype=np.float64)
    ls_ptr = lst_ptr = None
    if lists is not None:
        start = np.ascontiguousarray(lists[0], np.uint32)
        lst = np.ascontiguousarray(lists[1], np.uint32)
        ls_ptr, lst_ptr = _ptr(start), _ptr(lst)
    sg = np.empty((n, 9), np.float64) if splat_grad else None
    vis = lib().orc_grad_run(n, *[_ptr(a) for a in arrays], ctypes.byref(_cam(camera)), sh_degree, splat_scale,
                             tile[0], tile[1], ls_ptr, lst_ptr, _ptr(g), _ptr(image), _ptr(grad),
                             _ptr(sg) if splat_grad else None, threads)
    if splat_grad:
        return image, grad, vis, sg
    return image, grad, vis


def grad_fd(raw, camera, dLdC, idx, sh_degree=3, splat_scale=1.0, tile=(16, 16), h_rel=1e-6, threads=0):
    """FP64 central differences [m,59] of L = sum(dLdC * C) for Gaussians idx, and flags [m,59]
    (1: the perturbation changed the blend structure, 2: culled)."""
    arrays = _raw64(raw)
    idx = np.ascontiguousarray(idx, np.int64)
    fd = np.empty((len(idx), 59), np.float64)
    flags = np.empty((len(idx), 59), np.uint8)
    g = np.ascontiguousarray(dLdC, dtype=np.float64)
    lib().orc_grad_fd(raw.n, *[_ptr(a) for a in arrays], ctypes.byref(_cam(camera)), sh_degree, splat_scale, tile[0],
                      tile[1], _ptr(g), _ptr(idx), len(idx), h_rel, _ptr(fd), _ptr(flags), threads)
    return fd, flags
