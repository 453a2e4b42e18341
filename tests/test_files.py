"""COLMAP sparse models and PNG frames (aniso_amd.files, C ABI ans_colmap_load / ans_png_*), on
the CPU: against the compiled reference reader and writer (src/camera.cpp, src/image.cpp, built
into oracle/_ref when the reference sources are present) and, independently of it, against what
the file formats fix: zlib's own decoder, CRC-32 and Adler-32 for the PNG; scipy's quaternion
rotation and the written values for the camera files."""
import os
import struct
import zlib

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import oracle
from aniso_amd import files

needs_ref = pytest.mark.skipif(not oracle.ref_available(), reason="compiled reference not built")


def chunks(data: bytes):
    """(type, payload, crc ok) of every chunk after the signature."""
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    out, p = [], 8
    while p < len(data):
        n = struct.unpack(">I", data[p:p + 4])[0]
        kind, payload = data[p + 4:p + 8], data[p + 8:p + 8 + n]
        crc = struct.unpack(">I", data[p + 8 + n:p + 12 + n])[0]
        out.append((kind, payload, crc == zlib.crc32(kind + payload)))
        p += 12 + n
    assert p == len(data)
    return out


def decode(data: bytes):
    c = chunks(data)
    assert [k for k, _, _ in c] == [b"IHDR", b"IDAT", b"IEND"] and all(ok for _, _, ok in c)
    w, h, depth, colour, comp, filt, inter = struct.unpack(">IIBBBBB", c[0][1])
    assert (depth, colour, comp, filt, inter) == (8, 2, 0, 0, 0)
    raw = zlib.decompress(c[1][1])  # checks the stored blocks and the Adler-32
    assert struct.unpack(">I", c[1][1][-4:])[0] == zlib.adler32(raw)
    rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + 3 * w) if h else np.zeros((0, 1 + 3 * w), np.uint8)
    assert not rows[:, 0].any()  # filter type 0 on every scanline
    return rows[:, 1:].reshape(h, w, 3)


SIZES = [(3, 2), (1, 1), (0, 0), (0, 2), (1456, 15), (1456, 30), (1457, 15), (640, 360), (1920, 1080)]


def image(w, h, seed=0):
    return np.random.default_rng(seed).integers(0, 256, (h, w, 3), dtype=np.uint8)


def test_png_reference_size_check():
    """The reference's own check (tests/test_camera.cpp:58-64): a 3x2 frame with one red pixel is
    8 + 25 + (12 + 2 + 5 + 20 + 4) + 12 bytes."""
    img = np.zeros((2, 3, 3), np.uint8)
    img[1, 1, 0] = 255
    data = files.png_bytes(img)
    assert len(data) == 8 + 25 + (12 + 2 + 5 + 20 + 4) + 12
    assert np.array_equal(decode(data), img)


@pytest.mark.parametrize("w,h", SIZES)
def test_png_decodes_to_the_frame(w, h):
    """Stored blocks of at most 65535 bytes (1456x15 is exactly one full block, 1456x30 two, 1457x15
    one byte over), zlib stream, CRCs and Adler-32 all as a standard decoder reads them."""
    img = image(w, h, seed=w + h)
    data = files.png_bytes(img)
    assert np.array_equal(decode(data), img)
    raw_len = h * (1 + 3 * w)
    blocks = max(1, -(-raw_len // 65535))
    assert len(data) == 8 + 25 + 12 + (2 + 5 * blocks + raw_len + 4) + 12


@needs_ref
@pytest.mark.parametrize("w,h", SIZES)
def test_png_bytes_equal_reference_writer(tmp_path, w, h):
    img = image(w, h, seed=3 * w + h)
    ref = tmp_path / "ref.png"
    oracle.ref_write_png(ref, img)
    mine = tmp_path / "mine.png"
    files.write_png(mine, img)
    assert mine.read_bytes() == ref.read_bytes() == files.png_bytes(img)


def test_png_write_errors(tmp_path):
    with pytest.raises(ValueError, match="cannot write"):
        files.write_png(tmp_path / "missing" / "x.png", image(2, 2))
    with pytest.raises(ValueError):
        files.png_bytes(np.zeros((2, 2), np.uint8))


# ---- COLMAP -------------------------------------------------------------------------------------

def write_cameras(path, cams):
    """cams: (id, model, width, height, params)."""
    with open(path, "wb") as f:
        f.write(struct.pack("<Q", len(cams)))
        for cid, model, w, h, params in cams:
            f.write(struct.pack("<IiQQ", cid, model, w, h))
            f.write(struct.pack(f"<{len(params)}d", *params))


def write_images(path, images):
    """images: (image id, q (w, x, y, z), t, camera id, name, point count)."""
    with open(path, "wb") as f:
        f.write(struct.pack("<Q", len(images)))
        for iid, q, t, cid, name, npts in images:
            f.write(struct.pack("<I4d3dI", iid, *q, *t, cid))
            f.write(name.encode() + b"\0")
            f.write(struct.pack("<Q", npts))
            f.write(struct.pack("<ddq", 1.5, 2.5, -1) * npts)


def model(tmp_path, seed=0, n=6):
    rng = np.random.default_rng(seed)
    cams = [(1, 1, 1920, 1080, (1500.5, 1499.25, 960.1, 540.2)), (7, 0, 640, 480, (500.0, 320.5, 240.25))]
    images = []
    for k in range(n):
        q = rng.normal(size=4) * rng.uniform(0.5, 2.0)  # not unit: the reader normalizes
        t = rng.normal(size=3)
        images.append((100 + k, tuple(q), tuple(t), 1 if k % 2 else 7, f"img_{(k * 7) % n:03d}.jpg", k))
    write_cameras(tmp_path / "cameras.bin", cams)
    write_images(tmp_path / "images.bin", images)
    return cams, images


def test_colmap_values_and_order(tmp_path):
    cams, images = model(tmp_path)
    names, got = files.load_colmap(tmp_path)
    assert names == sorted(i[4] for i in images)
    by_name = {i[4]: i for i in images}
    intr = {c[0]: c for c in cams}
    for name, c in zip(names, got):
        iid, q, t, cid, _, _ = by_name[name]
        w, x, y, z = q
        R = Rotation.from_quat([x, y, z, w]).as_matrix()  # scipy's scalar-last convention
        assert np.allclose(c.R, R, atol=1e-12)
        assert np.array_equal(c.t, np.array(t))
        _, mdl, W, H, p = intr[cid]
        assert (c.width, c.height) == (W, H)
        if mdl == 0:
            assert (c.fx, c.fy, c.cx, c.cy) == (p[0], p[0], p[1], p[2])
        else:
            assert (c.fx, c.fy, c.cx, c.cy) == tuple(p)


@needs_ref
def test_colmap_equals_reference_reader(tmp_path):
    model(tmp_path, seed=4, n=9)
    names, got = files.load_colmap(tmp_path)
    ref_names, ref = oracle.ref_load_colmap(tmp_path)
    assert names == ref_names
    for a, b in zip(got, ref):
        assert np.array_equal(a.R, b["R"]) and np.array_equal(a.t, b["t"])
        assert (a.fx, a.fy, a.cx, a.cy, a.width, a.height) == (b["fx"], b["fy"], b["cx"], b["cy"], b["width"],
                                                                b["height"])


def error_cases(tmp_path):
    """(name, setup) pairs; each setup writes a broken model into its own folder."""
    def missing(d):
        pass

    def truncated_cameras(d):
        write_cameras(d / "cameras.bin", [(1, 1, 10, 10, (1.0, 1.0, 5.0, 5.0))])
        data = (d / "cameras.bin").read_bytes()
        (d / "cameras.bin").write_bytes(data[:-3])

    def distortion(d):
        write_cameras(d / "cameras.bin", [(1, 4, 10, 10, tuple(range(8)))])

    def unknown_model(d):
        write_cameras(d / "cameras.bin", [(1, 99, 10, 10, ())])

    def missing_camera(d):
        write_cameras(d / "cameras.bin", [(1, 1, 10, 10, (1.0, 1.0, 5.0, 5.0))])
        write_images(d / "images.bin", [(1, (1, 0, 0, 0), (0, 0, 0), 2, "a.png", 0)])

    def no_images(d):
        write_cameras(d / "cameras.bin", [(1, 1, 10, 10, (1.0, 1.0, 5.0, 5.0))])

    def truncated_name(d):
        write_cameras(d / "cameras.bin", [(1, 1, 10, 10, (1.0, 1.0, 5.0, 5.0))])
        write_images(d / "images.bin", [(1, (1, 0, 0, 0), (0, 0, 0), 1, "abc.png", 0)])
        data = (d / "images.bin").read_bytes()
        (d / "images.bin").write_bytes(data[:-12])

    cases = {"missing": (missing, "cameras.bin: cannot open"),
             "truncated_cameras": (truncated_cameras, "cameras.bin: unexpected end of file"),
             "distortion": (distortion, "cameras.bin: camera model 4 has lens distortion; undistort the images first"),
             "unknown_model": (unknown_model, "cameras.bin: unknown camera model 99"),
             "missing_camera": (missing_camera, "images.bin: image a.png refers to a missing camera"),
             "no_images": (no_images, "images.bin: cannot open"),
             "truncated_name": (truncated_name, "images.bin: unexpected end of file")}
    out = []
    for name, (setup, msg) in cases.items():
        d = tmp_path / name
        d.mkdir()
        setup(d)
        out.append((d, msg))
    return out


def test_colmap_errors(tmp_path):
    for d, msg in error_cases(tmp_path):
        with pytest.raises(ValueError) as e:
            files.load_colmap(d)
        assert str(e.value) == os.path.join(str(d), msg.split(":")[0]) + ":" + msg.split(":", 1)[1], str(e.value)
        if oracle.ref_available():
            with pytest.raises(oracle.RefPlyError) as r:
                oracle.ref_load_colmap(d)
            assert str(r.value) == str(e.value)


def test_colmap_last_image_points_past_end(tmp_path):
    """The point list is skipped with a seek: a last image whose points run past the end of the
    file still loads (as in the reference, which reads nothing after it)."""
    write_cameras(tmp_path / "cameras.bin", [(1, 1, 10, 10, (1.0, 1.0, 5.0, 5.0))])
    write_images(tmp_path / "images.bin", [(1, (1, 0, 0, 0), (0, 0, 0), 1, "a.png", 3)])
    data = (tmp_path / "images.bin").read_bytes()
    (tmp_path / "images.bin").write_bytes(data[:-30])
    names, cams = files.load_colmap(tmp_path)
    assert names == ["a.png"]
    if oracle.ref_available():
        assert oracle.ref_load_colmap(tmp_path)[0] == ["a.png"]
