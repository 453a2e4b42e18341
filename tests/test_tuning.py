"""The persistent tuning cache (aniso_amd.tuning) on the CPU, with a stand-in renderer whose
autotune calls are counted: a miss measures and stores, a hit in the same or a fresh process reads
the file, keys separate devices, scenes, sizes and base options, a corrupt or foreign-version file
reads as empty."""
import json

import pytest

from aniso_amd import scenes, tuning
from aniso_amd.renderer import RenderOptions


class FakeRenderer:
    def __init__(self, n=1000, degree=3):
        self.n, self.sh_degree = n, degree
        self.calls = 0

    def autotune(self, camera, frame, base=None, repeats=3):
        self.calls += 1
        b = base or RenderOptions()
        return RenderOptions(b.sh_degree, b.splat_scale, (32, 16), 128, b.cull, b.graph), 0.25

    def autotune_views(self, cameras, outputs, base=None, repeats=3):
        self.calls += 1
        b = base or RenderOptions()
        return RenderOptions(b.sh_degree, b.splat_scale, (8, 8), 64, b.cull, b.graph), 1.5


@pytest.fixture
def cache(tmp_path, monkeypatch):
    path = tmp_path / "t" / "tuning.json"
    monkeypatch.setenv("ANISO_AMD_TUNING", str(path))
    monkeypatch.setattr(tuning, "_device", lambda: "fake-gpu|gfx950|256")
    return path


def test_miss_then_hit(cache):
    r = FakeRenderer()
    cam = scenes.pinhole_camera(640, 360)
    a = tuning.tuned(r, cam, None)
    assert r.calls == 1 and a.tile == (32, 16) and a.batch_depth == 128
    b = tuning.tuned(FakeRenderer(), cam, None)  # a new renderer, same scene size: read from the file
    assert b.tile == (32, 16) and b.batch_depth == 128
    entries = json.loads(cache.read_text())["entries"]
    assert len(entries) == 1 and next(iter(entries.values()))["ms"] == 0.25


def test_keys_separate(cache, monkeypatch):
    r = FakeRenderer()
    cams = [scenes.pinhole_camera(640, 360), scenes.pinhole_camera(1280, 720)]
    tuning.tuned(r, cams[0], None)
    tuning.tuned(r, cams[1], None)                                   # another size
    tuning.tuned(FakeRenderer(n=2000), cams[0], None)                # another scene
    tuning.tuned(r, cams[0], None, base=RenderOptions(sh_degree=1))  # another base
    tuning.tuned_views(r, cams, [None, None])                        # a batch
    tuning.tuned_views(r, cams[::-1], [None, None])                  # the batch in another order
    monkeypatch.setattr(tuning, "_device", lambda: "other-gpu|gfx942|304")
    tuning.tuned(r, cams[0], None)                                   # another device
    assert len(json.loads(cache.read_text())["entries"]) == 7
    before = r.calls
    tuning.tuned(r, cams[0], None)  # the other device's entry, now cached
    assert r.calls == before


def test_base_options_kept(cache):
    r = FakeRenderer()
    base = RenderOptions(sh_degree=2, splat_scale=0.5, cull=False, graph=False)
    tuning.tuned_views(r, [scenes.pinhole_camera(64, 64)], [None], base=base)
    o = tuning.tuned_views(r, [scenes.pinhole_camera(64, 64)], [None], base=base)
    assert r.calls == 1
    assert (o.sh_degree, o.splat_scale, o.cull, o.graph, o.tile, o.batch_depth) == (2, 0.5, False, False, (8, 8), 64)


@pytest.mark.parametrize("content", ["{not json", json.dumps({"version": 999, "entries": {"x": {}}}), "[]"])
def test_unreadable_file_is_empty(cache, content):
    cache.parent.mkdir(parents=True, exist_ok=True)
    cache.write_text(content)
    r = FakeRenderer()
    tuning.tuned(r, scenes.pinhole_camera(64, 64), None)
    assert r.calls == 1
    assert json.loads(cache.read_text())["version"] == tuning.VERSION
