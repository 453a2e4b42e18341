"""Writes tests/golden/golden.json: SHA-256 (first 16 hex digits) of the raw RGB8 images the
reference CPU renderer produces, compiled from its own sources into oracle/_ref by
`make -C oracle ref`, on synthetic scenes and the random scene of tests/test_gpu_render.cpp.

The scene-file and C1-C3 image hashes marked "survey" were recorded when the reference was first
built from source in this project's container; this script re-derives every image hash from
the compiled reference and checks it agrees with those.

    make -C oracle all ref && python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import numpy as np  # noqa: E402

import oracle  # noqa: E402
from aniso_amd import scenes  # noqa: E402

SURVEY = {
    "scene:synthetic-200k-640x360": "7d95aeff470bebc7",
    "scene:synthetic-741883-1280x720": "0b06bc044b0c6429",
    "image:synthetic-200k-640x360:sh3": "99b2a5aecfb61cc6",
    "image:synthetic-200k-640x360:sh0": "3ae294ee53597df1",
    "image:synthetic-741883-1280x720:sh3": "9bbadc4190b07e50",
    "image:clustered-2.2M:1280x720:sh3": "68439357b68e6021",
    "image:clustered-2.2M:1920x1080:sh3": "3fa52708f10a2dca",
}


def h(b: bytes) -> str:
    return hashlib.sha256(b).hexdigest()[:16]


def posed_cameras():
    """Views with rotation and translation (poseCamera, main.cpp:179-202)."""
    return {
        "yaw20-640x360": scenes.pose_camera([0.3, -0.2, 0.5], [0.34, 0.0, 0.94], [0, -1, 0], 640, 360, 60.0),
        "tilt-333x217": scenes.pose_camera([-0.5, 0.4, -1.0], [-0.1, 0.2, 1.0], [0.05, -1, 0], 333, 217, 70.0),
    }


def main():
    assert oracle.ref_available(), "build the compiled reference first: make -C oracle ref"
    out = {"_source": "compiled reference CPU renderer (src/render.cpp, sh.cpp, camera.cpp)", "hashes": {}}
    H = out["hashes"]
    # Scene files.
    for name in ("synthetic-200k-640x360", "synthetic-741883-1280x720"):
        kw = scenes.WORKLOADS[name]
        H["scene:" + name] = h(scenes.ply_bytes(scenes.synthetic_scene(kw["n"], kw["width"], kw["height"])))
    # Synthetic scenes, identity camera.
    c1 = oracle.activate(scenes.synthetic_scene(200_000, 640, 360))
    cam = scenes.pinhole_camera(640, 360)
    for deg in (3, 0):
        H[f"image:synthetic-200k-640x360:sh{deg}"] = h(oracle.ref_render(c1, cam, sh_degree=deg)[0].tobytes())
    for vname, vcam in posed_cameras().items():
        H[f"image:synthetic-200k:{vname}:sh3"] = h(oracle.ref_render(c1, vcam)[0].tobytes())
    H["image:synthetic-200k-640x360:splat0.5"] = h(oracle.ref_render(c1, cam, splat_scale=0.5)[0].tobytes())
    c2 = oracle.activate(scenes.synthetic_scene(741_883, 1280, 720))
    H["image:synthetic-741883-1280x720:sh3"] = h(oracle.ref_render(c2, scenes.pinhole_camera(1280, 720))[0].tobytes())
    c3 = oracle.activate(scenes.synthetic_scene(2_200_000, 1280, 720, clustered=True))
    for w, hh in ((1280, 720), (1920, 1080)):
        H[f"image:clustered-2.2M:{w}x{hh}:sh3"] = h(oracle.ref_render(c3, scenes.pinhole_camera(w, hh))[0].tobytes())
    # The reference GPU test's scene and cameras (test_gpu_render.cpp:41-58 for the scene, :21-27 and :69-70 for the cameras).
    mt = oracle.mt19937_scene(20000, 7)
    for w, hh in ((320, 240), (333, 217)):
        c = scenes.Camera(np.eye(3), np.zeros(3), 0.8 * w, 0.8 * w, w / 2.0, hh / 2.0, w, hh)
        for deg in (3, 0):
            img, st = oracle.ref_render(mt, c, sh_degree=deg)
            H[f"image:mt19937-20k:{w}x{hh}:sh{deg}"] = h(img.tobytes())
            H[f"stats:mt19937-20k:{w}x{hh}:sh{deg}"] = [int(st.visible), int(st.pairs)]
    for k, v in SURVEY.items():
        assert H[k] == v, (k, H[k], v)
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
