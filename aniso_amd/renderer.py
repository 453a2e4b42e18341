"""Python mirror of the reference's GPU renderer API (include/aniso/gpu_render.h:22-38,
include/aniso/render.h:9-36) over the C ABI:

    renderer = GpuRenderer(scene)                 # uploads once
    image = renderer.render(camera, stats, options)  # uint8 [H, W, 3], rows top to bottom
    GpuRenderer.available()

Errors raise (ValueError for bad arguments, RuntimeError for HIP failures), as the reference
throws std::runtime_error.
"""
from __future__ import annotations

import ctypes
import dataclasses
import os

import numpy as np

from . import _lib
from .scenes import Camera, RawScene


@dataclasses.dataclass
class RenderOptions:
    """aniso::RenderOptions (render.h:28-33) plus tile shape and LDS batch depth."""

    sh_degree: int = 3
    splat_scale: float = 1.0
    tile: tuple = (16, 16)
    batch_depth: int = 256
    cull: bool = True  # conservative per-8x8-block culling in the blend (output identical either way)
    graph: bool = True  # batched views: replay the captured HIP graph (else direct launches)

    def c(self) -> _lib.Options_:
        flags = (0 if self.cull else _lib.FLAG_NO_CULL) | (0 if self.graph else _lib.FLAG_NO_GRAPH)
        return _lib.Options_(int(self.sh_degree), float(self.splat_scale), int(self.tile[0]), int(self.tile[1]),
                             int(self.batch_depth), flags)

    @classmethod
    def from_c(cls, o: _lib.Options_) -> "RenderOptions":
        return cls(o.sh_degree, o.splat_scale, (o.tile_w, o.tile_h), o.batch_depth, not (o.flags & _lib.FLAG_NO_CULL),
                   not (o.flags & _lib.FLAG_NO_GRAPH))


@dataclasses.dataclass
class RenderStats:
    """aniso::RenderStats (render.h:9-13); GPU stage times from clock readings the frame's kernels take."""

    visible: int = 0
    tile_pairs: int = 0
    project_ms: float = 0.0
    sort_ms: float = 0.0
    blend_ms: float = 0.0
    pair_capacity: int = 0
    retries: int = 0


@dataclasses.dataclass
class PathStats:
    """ans_path_stats: frames handed to the sink, batched launches, their GPU time, wall time."""

    frames: int = 0
    chunks: int = 0
    gpu_ms: float = 0.0
    total_ms: float = 0.0


def camera_c(camera: Camera) -> _lib.Camera_:
    c = _lib.Camera_()
    for i, v in enumerate(np.asarray(camera.R, np.float64).reshape(9)):
        c.R[i] = float(v)
    for i, v in enumerate(np.asarray(camera.t, np.float64).reshape(3)):
        c.t[i] = float(v)
    c.fx, c.fy, c.cx, c.cy = float(camera.fx), float(camera.fy), float(camera.cx), float(camera.cy)
    c.width, c.height = int(camera.width), int(camera.height)
    return c


def _torch_stream():
    """torch's current HIP stream if torch is loaded with a GPU, else None (the default stream)."""
    import sys
    torch = sys.modules.get("torch")
    if torch is not None and torch.cuda.is_available():
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    return None


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class GpuRenderer:
    """Owns a device copy of one scene; renders it from any camera."""

    def __init__(self, scene, activated: bool = False):
        """`scene` is a RawScene (raw 3DGS values, activated on the host like the reference
        loader), or, with activated=True, any object with pos/scale/rot/opacity/sh arrays."""
        L = _lib.load()
        h = ctypes.c_void_p()
        if isinstance(scene, RawScene) and not activated:
            arrays = [_f32(scene.means), _f32(scene.log_scales), _f32(scene.quats), _f32(scene.opacity_logits),
                      _f32(scene.sh)]
            fn = L.ans_renderer_create_raw
        else:
            arrays = [_f32(scene.pos), _f32(scene.scale), _f32(scene.rot), _f32(scene.opacity), _f32(scene.sh)]
            fn = L.ans_renderer_create
        n = arrays[0].shape[0]
        degree = int(round(np.sqrt(arrays[4].shape[1]))) - 1 if n else 0
        _lib.check(fn(n, degree, *[_ptr(a) for a in arrays], ctypes.byref(h)))
        self._h = h
        self.n = n
        self.sh_degree = degree

    @classmethod
    def from_ply(cls, source) -> "GpuRenderer":
        """A renderer from a 3DGS PLY: a path (ans_renderer_load_ply, streamed from the file) or
        the file's bytes (ans_renderer_create_ply). The header is read on the host with the
        reference loader's rules; activation and the f_rest reorder run on the device."""
        L = _lib.load()
        h = ctypes.c_void_p()
        if isinstance(source, (bytes, bytearray, memoryview)):
            buf = np.frombuffer(source, np.uint8)
            _lib.check(L.ans_renderer_create_ply(_ptr(buf), buf.size, ctypes.byref(h)))
        else:
            _lib.check(L.ans_renderer_load_ply(os.fsencode(source), ctypes.byref(h)))
        self = cls.__new__(cls)
        self._h = h
        n, degree = ctypes.c_int64(), ctypes.c_int32()
        _lib.check(L.ans_renderer_scene(h, ctypes.byref(n), ctypes.byref(degree), None, None, None, None, None))
        self.n, self.sh_degree = n.value, degree.value
        return self

    def scene_arrays(self):
        """The device scene copied back, activated: dict of pos [n,3], scale [n,3], rot [n,4]
        (w, x, y, z), opacity [n], sh [n, (degree+1)^2, 3]."""
        k = (self.sh_degree + 1) ** 2
        out = dict(pos=np.empty((self.n, 3), np.float32), scale=np.empty((self.n, 3), np.float32),
                   rot=np.empty((self.n, 4), np.float32), opacity=np.empty(self.n, np.float32),
                   sh=np.empty((self.n, k, 3), np.float32))
        _lib.check(_lib.load().ans_renderer_scene(self._h, None, None, *[_ptr(out[key]) for key in
                                                                          ("pos", "scale", "rot", "opacity", "sh")]))
        return out

    @staticmethod
    def available() -> bool:
        return bool(_lib.load().ans_available())

    def render(self, camera: Camera, stats: RenderStats | None = None,
               options: RenderOptions | None = None) -> np.ndarray:
        img, _ = self.render_float(camera, stats, options, want_float=False)
        return img

    def render_float(self, camera: Camera, stats: RenderStats | None = None, options: RenderOptions | None = None,
                     want_float: bool = True):
        """(uint8 [H,W,3], float32 [H,W,3] unclamped colour or None)."""
        opts = (options or RenderOptions()).c()
        img = np.empty((camera.height, camera.width, 3), np.uint8)
        imgf = np.empty((camera.height, camera.width, 3), np.float32) if want_float else None
        st = _lib.Stats_()
        _lib.check(_lib.load().ans_render_float(self._h, ctypes.byref(camera_c(camera)), ctypes.byref(opts),
                                                _ptr(img), _ptr(imgf) if want_float else None,
                                                camera.width * camera.height, ctypes.byref(st)))
        if stats is not None:
            for f in dataclasses.fields(RenderStats):
                setattr(stats, f.name, getattr(st, f.name))
        return img, imgf

    def render_views(self, cameras, outputs, options: RenderOptions | None = None, stats: list | None = None):
        """Renders every camera in one batched pass (ans_render_views) on torch's current stream.
        outputs: one device buffer per camera (torch uint8 tensors of H*W*3 elements on the GPU,
        or raw device addresses). Fills `stats` (a list) with one RenderStats per view if given."""
        k = len(cameras)
        if len(outputs) != k:
            raise ValueError("one output per camera")
        cams = (_lib.Camera_ * k)(*[camera_c(c) for c in cameras])
        ptrs = (ctypes.c_void_p * k)(*[o if isinstance(o, int) else o.data_ptr() for o in outputs])
        st = (_lib.Stats_ * k)()
        opts = (options or RenderOptions()).c()
        _lib.check(_lib.load().ans_render_views(self._h, k, cams, ctypes.byref(opts), ptrs, _torch_stream(), st))
        if stats is not None:
            stats.clear()
            for s_ in st:
                stats.append(RenderStats(**{f.name: getattr(s_, f.name) for f in dataclasses.fields(RenderStats)}))

    def render_views_async(self, cameras, outputs, options: RenderOptions | None = None):
        """Enqueues one batched pass (ans_render_views_async) on torch's current stream and returns;
        views_wait() waits for every batch enqueued so. The output buffers must stay alive until
        then."""
        k = len(cameras)
        if len(outputs) != k:
            raise ValueError("one output per camera")
        cams = (_lib.Camera_ * k)(*[camera_c(c) for c in cameras])
        ptrs = (ctypes.c_void_p * k)(*[o if isinstance(o, int) else o.data_ptr() for o in outputs])
        opts = (options or RenderOptions()).c()
        _lib.check(_lib.load().ans_render_views_async(self._h, k, cams, ctypes.byref(opts), ptrs, _torch_stream()))
        self._async_views = k

    def views_wait(self, stats: list | None = None) -> bool:
        """Waits for every batch of render_views_async. Returns False if a view of any of them had
        more pairs than the pair arena held (ANS_ERR_CAPACITY: the arena has grown; render those
        batches again), True otherwise. Fills `stats` (a list) with the last batch's RenderStats."""
        n_views = getattr(self, "_async_views", 0)
        st = (_lib.Stats_ * max(1, n_views))()
        rc = _lib.load().ans_views_wait(self._h, st if n_views > 0 else None)
        if rc == 3:  # ANS_ERR_CAPACITY
            return False
        _lib.check(rc)
        if stats is not None and n_views > 0:
            stats.clear()
            for s_ in st:
                stats.append(RenderStats(**{f.name: getattr(s_, f.name) for f in dataclasses.fields(RenderStats)}))
        return True

    def autotune(self, camera: Camera, frame, base: RenderOptions | None = None, repeats: int = 3):
        """(best RenderOptions, device ms per frame) over tile shapes x LDS batch depths for single
        frames of `camera`, rendered into the device buffer `frame` (ans_render_autotune)."""
        best = _lib.Options_()
        ms = ctypes.c_double()
        ptr = frame if isinstance(frame, int) else frame.data_ptr()
        _lib.check(_lib.load().ans_render_autotune(self._h, ctypes.byref(camera_c(camera)),
                                                   ctypes.byref((base or RenderOptions()).c()), ctypes.c_void_p(ptr),
                                                   int(repeats), ctypes.byref(best), ctypes.byref(ms)))
        return RenderOptions.from_c(best), ms.value

    def autotune_views(self, cameras, outputs, base: RenderOptions | None = None, repeats: int = 3):
        """(best RenderOptions, batch ms) over tile shapes x LDS batch depths (ans_views_autotune)."""
        k = len(cameras)
        cams = (_lib.Camera_ * k)(*[camera_c(c) for c in cameras])
        ptrs = (ctypes.c_void_p * k)(*[o if isinstance(o, int) else o.data_ptr() for o in outputs])
        best = _lib.Options_()
        ms = ctypes.c_double()
        _lib.check(_lib.load().ans_views_autotune(self._h, k, cams, ctypes.byref((base or RenderOptions()).c()), ptrs,
                                                  _torch_stream(), int(repeats), ctypes.byref(best),
                                                  ctypes.byref(ms)))
        return RenderOptions.from_c(best), ms.value

    def render_path(self, cameras, sink, options: RenderOptions | None = None, chunk: int = 16,
                    splat_scales=None) -> "PathStats":
        """Renders a camera path (ans_render_path): chunks of `chunk` views in batched launches, two
        chunks in flight, every frame handed to `sink` in order as it reaches host memory.

        sink: a callable sink(frame_index, rgb) with rgb a uint8 [H, W, 3] array that is only
        valid during the call (copy it to keep it), or a binary file object (each frame is
        written, as `aniso path` writes raw RGB to stdout, src/main.cpp:243). A callable that
        returns False stops the path (ValueError); an exception it raises stops the path and is
        raised again here. splat_scales: None (options.splat_scale throughout) or one scale per
        frame, the reference path file's optional tenth column."""
        k = len(cameras)
        cams = (_lib.Camera_ * max(k, 1))(*[camera_c(c) for c in cameras])
        sizes = [(c.height, c.width) for c in cameras]
        scales = None
        if splat_scales is not None:
            scales = np.ascontiguousarray(splat_scales, np.float32)
            if scales.shape != (k,):
                raise ValueError("one splat scale per camera")
        write = getattr(sink, "write", None)
        failure = []

        def deliver(_user, frame, rgb, nbytes):
            try:
                buf = np.ctypeslib.as_array(rgb, shape=(nbytes,))
                if write is not None:
                    write(buf.tobytes())
                    return 0
                h, w = sizes[frame]
                return 1 if sink(frame, buf.reshape(h, w, 3)) is False else 0
            except BaseException as e:  # handed back to the caller after the library drains
                failure.append(e)
                return 1

        cb = _lib.FrameSink(deliver)
        st = _lib.PathStats_()
        opts = (options or RenderOptions()).c()
        rc = _lib.load().ans_render_path(self._h, k, cams, None if scales is None else _ptr(scales),
                                         ctypes.byref(opts), int(chunk), cb, None, ctypes.byref(st))
        if failure:
            raise failure[0]
        _lib.check(rc)
        return PathStats(st.frames, st.chunks, st.gpu_ms, st.total_ms)

    def views_dump(self, view: int):
        """Lists of one view of the last batched render: dict with visible, pairs, values u32[P],
        ranges u32[T,2]."""
        L = _lib.load()
        vals = [ctypes.c_int64() for _ in range(3)]
        _lib.check(L.ans_views_frame_counts(self._h, view, *[ctypes.byref(v) for v in vals]))
        visible, pairs, tiles = (v.value for v in vals)
        out = dict(visible=visible, pairs=pairs, values=np.empty(pairs, np.uint32),
                   ranges=np.empty((tiles, 2), np.uint32))
        _lib.check(L.ans_views_frame_dump(self._h, view, _ptr(out["values"]), _ptr(out["ranges"])))
        return out

    def frame_counts(self):
        vals = [ctypes.c_int64() for _ in range(4)]
        _lib.check(_lib.load().ans_frame_counts(self._h, *[ctypes.byref(v) for v in vals]))
        return tuple(v.value for v in vals)  # n, visible, pairs, tiles

    def dump(self):
        """Intermediates of the last frame: dict with splat_bits u32[n,10], rect i32[n,4],
        keys u64[P], values u32[P], ranges u32[T,2]."""
        n, _, pairs, tiles = self.frame_counts()
        out = dict(splat_bits=np.empty((n, 10), np.uint32), rect=np.empty((n, 4), np.int32),
                   keys=np.empty(pairs, np.uint64), values=np.empty(pairs, np.uint32),
                   ranges=np.empty((tiles, 2), np.uint32))
        _lib.check(_lib.load().ans_frame_dump(self._h, *[_ptr(out[k]) for k in
                                                          ("splat_bits", "rect", "keys", "values", "ranges")]))
        return out

    def close(self):
        if getattr(self, "_h", None):
            _lib.load().ans_renderer_destroy(self._h)
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


def ply_inspect(data) -> tuple:
    """(Gaussian count, SH degree) of a PLY file image, from its header alone (ans_ply_inspect;
    no GPU). Malformed headers raise ValueError with the reference loader's message."""
    buf = np.frombuffer(data, np.uint8)
    n, degree = ctypes.c_int64(), ctypes.c_int32()
    _lib.check(_lib.load().ans_ply_inspect(_ptr(buf), buf.size, ctypes.byref(n), ctypes.byref(degree)))
    return n.value, degree.value


def activate(scene: RawScene):
    """Host activation as the renderer library does it (src/ply_loader.cpp:124-139); returns
    (scale [n,3], opacity [n], rot [n,4])."""
    n = scene.n
    scale = np.empty((n, 3), np.float32)
    opacity = np.empty(n, np.float32)
    rot = np.empty((n, 4), np.float32)
    _lib.load().ans_activate(n, _ptr(_f32(scene.log_scales)), _ptr(_f32(scene.opacity_logits)),
                             _ptr(_f32(scene.quats)), _ptr(scale), _ptr(opacity), _ptr(rot))
    return scale, opacity, rot
