"""The header-only C++ drop-in (aniso_amd/gpu_render.hpp) compiles against the reference's own
headers and links against libaniso_amd.so, with every call the reference's callers make
(constructor, render with stats and options, available) and the path-streaming addition. Runs where the reference tree is present
(it is read for its headers only) and g++ exists; it does not touch a GPU."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_INCLUDE = "/root/reference/include"

PROGRAM = r"""
#include "gpu_render.hpp"
#include <cstdio>
int main(int argc, char**) {
    if (argc > 5) {  // never taken: makes the program link every call of the shim
        aniso::Scene scene;
        aniso::GpuRenderer renderer(scene);
        aniso::Camera camera;
        aniso::RenderStats stats;
        aniso::RenderOptions options;
        options.shDegree = 1;
        aniso::Image image = renderer.render(camera, &stats, options);
        std::printf("%d\n", image.width);
        std::vector<aniso::Camera> path(3, camera);
        renderer.renderPath(path, {}, options, [](std::size_t, const uint8_t* rgb, std::size_t bytes) {
            std::fwrite(rgb, 1, bytes, stdout);
        });
    }
    std::printf("available=%d\n", aniso::GpuRenderer::available() ? 1 : 0);
    return 0;
}
"""


@pytest.mark.skipif(not os.path.isdir(REF_INCLUDE) or shutil.which("g++") is None,
                    reason="needs the reference headers and g++")
def test_shim_compiles_links_and_runs(tmp_path):
    src = tmp_path / "shim_main.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "shim_main"
    lib_dir = os.path.join(ROOT, "aniso_amd")
    cmd = ["g++", "-std=c++20", "-O1", "-Wall", "-Werror", f"-I{REF_INCLUDE}", f"-I{ROOT}/include", f"-I{lib_dir}",
           str(src), f"-L{lib_dir}", "-laniso_amd", f"-Wl,-rpath,{lib_dir}", "-Wl,-rpath-link,/opt/rocm/lib",
           "-o", str(exe)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, run.stderr
    assert run.stdout.startswith("available=")
