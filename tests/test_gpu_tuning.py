"""Autotuning on the GPU: ans_render_autotune picks one of the nine supported (tile, depth)
configurations for single frames, the chosen options render a close image to the defaults', and
the tuning cache reuses the answer from its file."""
import json

import numpy as np
import pytest
import torch

from aniso_amd import scenes, tuning
from aniso_amd.renderer import GpuRenderer, RenderOptions

pytestmark = pytest.mark.gpu


def test_single_frame_autotune_and_cache(gpu, tmp_path, monkeypatch):
    monkeypatch.setenv("ANISO_AMD_TUNING", str(tmp_path / "tuning.json"))
    raw = scenes.synthetic_scene(100_000, 640, 360, seed=2)
    cam = scenes.pinhole_camera(640, 360)
    frame = torch.empty(640 * 360 * 3, dtype=torch.uint8, device="cuda")
    with GpuRenderer(raw) as r:
        best, ms = r.autotune(cam, frame, repeats=2)
        print(f"case autotune 100k 640x360: {best.tile} depth {best.batch_depth} {ms:.4f} ms", flush=True)
        assert best.tile in [(8, 8), (16, 16), (32, 16)] and best.batch_depth in (64, 128, 256)
        assert 0 < ms < 100
        tuned = tuning.tuned(r, cam, frame)
        entries = json.loads((tmp_path / "tuning.json").read_text())["entries"]
        assert len(entries) == 1
        again = tuning.tuned(r, cam, frame)
        assert (again.tile, again.batch_depth) == (tuned.tile, tuned.batch_depth)
        # Other tile shapes change which splats a tile lists (the 3-sigma rect decides), so the
        # image may differ from the default one where a splat reaches past its rect.
        a = r.render(cam, options=tuned).astype(np.float64)
        b = r.render(cam).astype(np.float64)
        mse = np.mean((a - b) ** 2)
        assert mse == 0 or 10 * np.log10(255.0 ** 2 / mse) > 35
