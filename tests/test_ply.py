"""PLY ingest on the host side (no GPU): the header rules of ans_ply_inspect against the
reference loader (aniso::loadPly, src/ply_loader.cpp, compiled from its own sources into
oracle/_ref), the known answers of the reference's own loader test (tests/test_ply_loader.cpp),
and the device activation's expf restatement against the C library's expf.
"""
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
from aniso_amd import scenes
from aniso_amd.renderer import ply_inspect

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def ply_file(props, rows, fmt="binary_little_endian", extra_header=(), newline="\n", elements=()):
    """A PLY file image: vertex properties `props` (all float), rows [n, len(props)] float32."""
    rows = np.asarray(rows, np.float32).reshape(-1, len(props))
    lines = ["ply", f"format {fmt} 1.0", *extra_header, f"element vertex {rows.shape[0]}"]
    lines += [f"property float {p}" for p in props]
    for name, count, eprops in elements:
        lines.append(f"element {name} {count}")
        lines += [f"property {p}" for p in eprops]
    lines.append("end_header")
    return (newline.join(lines) + newline).encode() + rows.astype("<f4").tobytes()


def reference_test_scene(degree, count):
    """The scene tests/test_ply_loader.cpp:28-53 writes: Gaussian i has position (1+i, 2+i, 3+i),
    zero normals, f_dc (0.1, 0.2, 0.3), f_rest_k = 100 + k, opacity logit 0, log-scales
    (0, ln 2, ln 0.5), quaternion (2, 0, 0, 0)."""
    rest = 3 * ((degree + 1) ** 2 - 1)
    props = ["x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"]
    props += [f"f_rest_{k}" for k in range(rest)]
    props += ["opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"]
    rows = []
    for i in range(count):
        rows.append([1.0 + i, 2.0 + i, 3.0 + i, 0, 0, 0, 0.1, 0.2, 0.3] + [100.0 + k for k in range(rest)]
                    + [0.0, 0.0, math.log(2.0), math.log(0.5), 2.0, 0.0, 0.0, 0.0])
    return ply_file(props, rows)


@pytest.mark.parametrize("degree", [0, 1, 3])
def test_inspect_reference_test_scenes(degree):
    assert ply_inspect(reference_test_scene(degree, 3)) == (3, degree)


def standard_props(degree=3):
    return (["x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"]
            + [f"f_rest_{k}" for k in range(3 * ((degree + 1) ** 2 - 1))]
            + ["opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"])


def malformed_files():
    P = standard_props(1)
    row = np.zeros((2, len(P)))
    good = ply_file(P, row)
    return {
        "not_ply": b"plx\nformat binary_little_endian 1.0\nend_header\n",
        "ascii": b"ply\nformat ascii 1.0\nelement vertex 0\nend_header\n",  # test_ply_loader.cpp:75-86
        "unexpected_element": ply_file(P, row, elements=[("face", 2, ["list uchar int vertex_indices"])]),
        "uchar_property": ply_file(P[:-1], row[:, :-1], extra_header=()).replace(
            b"property float rot_2\n", b"property uchar rot_2\n"),
        "list_property": good.replace(b"property float nx\n", b"property list uchar int nx\n"),
        "missing_format": good.replace(b"format binary_little_endian 1.0\n", b""),
        "no_end_header": b"ply\nformat binary_little_endian 1.0\nelement vertex 1\nproperty float x\n",
        "missing_rot_3": ply_file(P[:-1], row[:, :-1]),
        "bad_rest_count": ply_file(P + ["f_rest_9"], np.zeros((2, len(P) + 1))),
        "short_body": good[:-5],
    }


def valid_variants():
    P = standard_props(2)
    rng = np.random.default_rng(3)
    rows = rng.normal(0, 1, (4, len(P)))
    shuffled = list(P)
    rng.shuffle(shuffled)
    return {
        "crlf": ply_file(P, rows, newline="\r\n"),
        "empty_other_element": ply_file(P, rows, elements=[("face", 0, ["list uchar int vertex_indices"])]),
        "comment_lines": ply_file(P, rows, extra_header=["comment made by a test", "obj_info x"]),
        "shuffled_and_extra": ply_file(shuffled + ["red", "x"], np.concatenate([rows, rows[:, :2]], 1)),
        "float32_type": ply_file(P, rows).replace(b"property float x\n", b"property float32 x\n"),
        "empty": ply_file(P, np.zeros((0, len(P)))),
    }


@pytest.mark.parametrize("case", sorted(malformed_files()))
def test_header_errors_match_reference_loader(case, tmp_path):
    data = malformed_files()[case]
    with pytest.raises(ValueError) as ours:
        ply_inspect(data)
    msg = str(ours.value)
    assert msg.startswith("PLY data: ")
    if oracle.ref_available():
        path = tmp_path / f"{case}.ply"
        path.write_bytes(data)
        with pytest.raises(oracle.RefPlyError) as ref:
            oracle.ref_load_ply(path)
        assert msg[len("PLY data: "):] == str(ref.value)[len(str(path)) + 2:]


@pytest.mark.parametrize("case", sorted(valid_variants()))
def test_valid_headers_match_reference_loader(case, tmp_path):
    data = valid_variants()[case]
    n, degree = ply_inspect(data)
    if oracle.ref_available():
        path = tmp_path / f"{case}.ply"
        path.write_bytes(data)
        ref, _ = oracle.ref_load_ply(path)
        assert (n, degree) == (ref["pos"].shape[0], ref["sh_degree"])


def test_inspect_synthetic_scene_file():
    sc = scenes.synthetic_scene(1000, 320, 240, seed=2)
    assert ply_inspect(scenes.ply_bytes(sc)) == (1000, 3)


@pytest.mark.skipif(shutil.which("/opt/rocm/bin/hipcc") is None, reason="needs hipcc")
def test_expf_restatement_matches_libm(tmp_path):
    """The device activation's expf (activate.hpp, compiled here for the host from the same source)
    equals the C library's expf, which the reference loader calls, on every 13th float32."""
    exe = tmp_path / "expf_check"
    cmd = ["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O2", "-ffp-contract=off",
           "-I" + os.path.join(ROOT, "aniso_amd", "csrc"), os.path.join(ROOT, "oracle", "expf_check.cpp"),
           "-o", str(exe)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    run = subprocess.run([str(exe), "13"], capture_output=True, text=True, timeout=300)
    assert run.returncode == 0, run.stdout
    assert run.stdout.strip().endswith("differ 0")
