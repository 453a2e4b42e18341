"""The C-ABI library builds, loads without a GPU, and exports every symbol include/aniso_amd.h
declares; its host-side activation equals the reference loader's (via the oracle)."""
import ctypes
import os
import re

import numpy as np

import oracle
from aniso_amd import _lib, renderer, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "aniso_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ans_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    names = declared_symbols()
    assert len(names) >= 10
    for name in names:
        assert hasattr(L, name), name
    assert set(names) == set(_lib.SIGNATURES), "binding table out of date with the header"


def test_last_error_is_empty_initially():
    assert _lib.load().ans_last_error() == b""


def test_create_rejects_bad_arguments_without_touching_the_device():
    L = _lib.load()
    h = ctypes.c_void_p()
    assert L.ans_renderer_create(-1, 3, None, None, None, None, None, ctypes.byref(h)) == 1
    assert b"count" in L.ans_last_error()
    assert L.ans_renderer_create(10, 7, None, None, None, None, None, ctypes.byref(h)) == 1


def test_host_activation_matches_reference_loader():
    raw = scenes.synthetic_scene(50_000, 640, 360, seed=11)
    raw.quats[:10] *= 3.7  # unnormalized quaternions
    raw.quats[10] = 0.0    # |q| = 0 gives the identity (ply_loader.cpp:130)
    scale, opacity, rot = renderer.activate(raw)
    act = oracle.activate(raw)
    assert np.array_equal(scale.view(np.uint32), act.scale.view(np.uint32))
    assert np.array_equal(opacity.view(np.uint32), act.opacity.view(np.uint32))
    assert np.array_equal(rot.view(np.uint32), act.rot.view(np.uint32))
    assert np.array_equal(rot[10], [1, 0, 0, 0])
