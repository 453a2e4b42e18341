"""Tuned render options that persist across processes.

The tile shape and LDS batch depth that render a scene fastest depend on the device, the scene
and the image sizes; GpuRenderer.autotune (single frames, ans_render_autotune) and
GpuRenderer.autotune_views (batches, ans_views_autotune) measure every combination. This module
keeps their answers in a small JSON file, keyed by device, scene size, SH degree and image sizes,
so that a later process reuses them instead of measuring again:

    opts = tuning.tuned(renderer, cam, frame)              # single frames of this camera size
    opts = tuning.tuned_views(renderer, cams, outputs)     # batches of these camera sizes

The file is ANISO_AMD_TUNING if set, else ~/.cache/aniso_amd/tuning.json. Writes go through a
temporary file and a rename, so concurrent processes never read a half-written file (the last
writer's entries win).
"""
from __future__ import annotations

import json
import os
import tempfile

from .renderer import RenderOptions

VERSION = 1


def cache_path() -> str:
    return os.environ.get("ANISO_AMD_TUNING") or os.path.join(os.path.expanduser("~"), ".cache", "aniso_amd",
                                                              "tuning.json")


def _device() -> str:
    import torch

    p = torch.cuda.get_device_properties(torch.cuda.current_device())
    return f"{p.name}|{getattr(p, 'gcnArchName', '')}|{p.multi_processor_count}"


def key(kind: str, renderer, sizes, base: RenderOptions | None = None) -> str:
    """The entry for `kind` ("frame" or "views") on this device and scene: the scene's Gaussian
    count and SH degree, the image sizes (in order: a batch's views are timed in order) and the
    options that are not tuned."""
    b = base or RenderOptions()
    n, degree = renderer.n, renderer.sh_degree
    dims = ",".join(f"{w}x{h}" for w, h in sizes)
    return f"{kind}|{_device()}|n={n}|sh={degree}|{dims}|deg={b.sh_degree}|scale={b.splat_scale:.6g}|cull={int(b.cull)}"


def load(path: str | None = None) -> dict:
    path = path or cache_path()
    try:
        with open(path) as f:
            data = json.load(f)
    except (OSError, ValueError):
        return {}
    if not isinstance(data, dict) or data.get("version") != VERSION or not isinstance(data.get("entries"), dict):
        return {}
    return data["entries"]


def store(entry: str, options: RenderOptions, ms: float, path: str | None = None) -> None:
    path = path or cache_path()
    entries = load(path)
    entries[entry] = {"tile": list(options.tile), "batch_depth": options.batch_depth, "ms": ms}
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    fd, tmp = tempfile.mkstemp(dir=os.path.dirname(os.path.abspath(path)), suffix=".tmp")
    try:
        with os.fdopen(fd, "w") as f:
            json.dump({"version": VERSION, "entries": entries}, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except BaseException:
        if os.path.exists(tmp):
            os.remove(tmp)
        raise


def _options(base: RenderOptions | None, hit: dict) -> RenderOptions:
    b = base or RenderOptions()
    return RenderOptions(b.sh_degree, b.splat_scale, tuple(hit["tile"]), int(hit["batch_depth"]), b.cull, b.graph)


def tuned(renderer, camera, frame, base: RenderOptions | None = None, path: str | None = None,
          repeats: int = 3) -> RenderOptions:
    """Options for single frames of `camera`'s size: the cached entry, or ans_render_autotune's
    answer (rendered into the device buffer `frame`), which is then cached."""
    entry = key("frame", renderer, [(camera.width, camera.height)], base)
    hit = load(path).get(entry)
    if hit:
        return _options(base, hit)
    best, ms = renderer.autotune(camera, frame, base, repeats)
    store(entry, best, ms, path)
    return best


def tuned_views(renderer, cameras, outputs, base: RenderOptions | None = None, path: str | None = None,
                repeats: int = 3) -> RenderOptions:
    """Options for batches of these cameras' sizes (in order): cached, or ans_views_autotune's."""
    entry = key("views", renderer, [(c.width, c.height) for c in cameras], base)
    hit = load(path).get(entry)
    if hit:
        return _options(base, hit)
    best, ms = renderer.autotune_views(cameras, outputs, base, repeats)
    store(entry, best, ms, path)
    return best
