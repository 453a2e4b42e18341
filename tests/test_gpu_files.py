"""PNG file images encoded on the GPU (ans_png_encode_device through aniso_amd.files.png_device)
against the host encoder, whose bytes tests/test_files.py pins to the reference writer and to
zlib's decoder: every byte equal, including the stored-block headers at 65535-byte boundaries,
the Adler-32 and the CRC-32 combined from 256-byte pieces."""
import numpy as np
import pytest
import torch

from aniso_amd import files, scenes
from aniso_amd.renderer import GpuRenderer, RenderOptions

pytestmark = pytest.mark.gpu

SIZES = [(1, 1), (3, 2), (1456, 15), (1456, 30), (1457, 15), (640, 360), (1920, 1080), (4096, 2160)]


@pytest.mark.parametrize("w,h", SIZES)
def test_device_png_equals_host_png(gpu, w, h):
    print(f"case png {w}x{h}", flush=True)
    img = np.random.default_rng(w * 7 + h).integers(0, 256, (h, w, 3), dtype=np.uint8)
    dev = files.png_device(torch.from_numpy(img).cuda())
    torch.cuda.synchronize()
    assert dev.cpu().numpy().tobytes() == files.png_bytes(img)


def test_device_png_reuses_buffers_and_encodes_a_rendered_frame(gpu):
    """One output and workspace for two frames in a row (the second smaller): each file image is
    that frame's; then a rendered frame, still on the device, to its PNG."""
    big = np.random.default_rng(1).integers(0, 256, (720, 1280, 3), dtype=np.uint8)
    small = np.random.default_rng(2).integers(0, 256, (180, 320, 3), dtype=np.uint8)
    out = torch.empty(4 << 20, dtype=torch.uint8, device="cuda")
    work = torch.empty(1 << 20, dtype=torch.uint8, device="cuda")
    a = files.png_device(torch.from_numpy(big).cuda(), out, work).cpu().numpy().tobytes()
    b = files.png_device(torch.from_numpy(small).cuda(), out, work).cpu().numpy().tobytes()
    assert a == files.png_bytes(big) and b == files.png_bytes(small)
    raw = scenes.synthetic_scene(50_000, 640, 360, seed=3)
    cam = scenes.pinhole_camera(640, 360)
    frame = torch.empty(360 * 640 * 3, dtype=torch.uint8, device="cuda")
    with GpuRenderer(raw) as r:
        r.render_views([cam], [frame], RenderOptions())
        torch.cuda.synchronize()
        assert np.array_equal(frame.cpu().numpy().reshape(360, 640, 3), r.render(cam))
    host = frame.cpu().numpy().reshape(360, 640, 3)
    dev = files.png_device(frame.view(360, 640, 3))
    assert dev.cpu().numpy().tobytes() == files.png_bytes(host)
