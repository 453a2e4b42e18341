// Drop-in for the reference's aniso::GpuRenderer (include/aniso/gpu_render.h:22-38) over the
// libaniso_amd C ABI (include/aniso_amd.h). Header-only: a reference build that includes this
// file where it included "aniso/gpu_render.h" and links libaniso_amd.so instead of compiling
// src/gpu_render.cu keeps its callers unchanged (src/main.cpp:158-165 and :227-239,
// src/viewer.cpp:133 and :244, tests/test_gpu_render.cpp:68-89).
//
// Same contract as the reference class: the scene is uploaded once by the constructor
// (gpu_render.cu:267-292), render() returns a host RGB8 image with rows top to bottom
// (gpu_render.cu:352-373), stats carry GPU stage times, errors throw std::runtime_error
// (gpu_render.cu:18-21), the class is neither copyable nor assignable. Differences: tiles are
// 16x16 (the reference's too, gpu_render.cu:13); stats->visible holds the visible count (the
// reference leaves it 0 on the GPU path).
#pragma once

#include "aniso/camera.h"
#include "aniso/gaussian.h"
#include "aniso/image.h"
#include "aniso/render.h"
#include "aniso_amd.h"

#include <cstddef>
#include <exception>
#include <stdexcept>
#include <string>
#include <vector>

namespace aniso {

class GpuRenderer {
public:
    explicit GpuRenderer(const Scene& scene) {
        const std::size_t n = scene.size();
        std::vector<float> pos(3 * n), scale(3 * n), rot(4 * n), opacity(n);
        for (std::size_t i = 0; i < n; ++i) {
            const Gaussian& g = scene.gaussians[i];
            pos[3 * i + 0] = g.position.x;
            pos[3 * i + 1] = g.position.y;
            pos[3 * i + 2] = g.position.z;
            scale[3 * i + 0] = g.scale.x;
            scale[3 * i + 1] = g.scale.y;
            scale[3 * i + 2] = g.scale.z;
            rot[4 * i + 0] = g.rotation.w;
            rot[4 * i + 1] = g.rotation.x;
            rot[4 * i + 2] = g.rotation.y;
            rot[4 * i + 3] = g.rotation.z;
            opacity[i] = g.opacity;
        }
        check(ans_renderer_create(static_cast<int64_t>(n), scene.shDegree, pos.data(), scale.data(), rot.data(),
                                  opacity.data(), scene.sh.data(), &handle_));
    }
    ~GpuRenderer() { ans_renderer_destroy(handle_); }
    GpuRenderer(const GpuRenderer&) = delete;
    GpuRenderer& operator=(const GpuRenderer&) = delete;

    // Timings in stats are GPU time per stage, from clock readings the frame's own kernels take.
    Image render(const Camera& camera, RenderStats* stats = nullptr, const RenderOptions& options = {}) {
        const ans_camera c = toC(camera);
        ans_options o{};
        o.sh_degree = options.shDegree;
        o.splat_scale = options.splatScale;  // tile shape and batch depth: the library defaults
        Image image(camera.width, camera.height);
        ans_stats st{};
        check(ans_render(handle_, &c, &o, image.rgb.data(), image.rgb.size(), &st));
        if (stats) {
            stats->visible = static_cast<std::size_t>(st.visible);
            stats->tilePairs = static_cast<std::size_t>(st.tile_pairs);
            stats->projectMs = st.project_ms;
            stats->sortMs = st.sort_ms;
            stats->blendMs = st.blend_ms;
        }
        return image;
    }

    // True if a HIP device is present and usable.
    static bool available() { return ans_available() != 0; }

    // Camera-path streaming, an addition to the reference class for the frame loop of
    // `aniso path` (src/main.cpp:234-244): renders the cameras in batched chunks, two chunks in
    // flight, and calls sink(index, rgb, bytes) for every frame in order, rgb pointing at RGB8
    // rows top to bottom that stay valid until sink returns. splatScales: empty (the options'
    // scale throughout) or one per camera. A sink that throws stops the path; the exception
    // reaches the caller after the GPU work has drained.
    template <class Sink>
    void renderPath(const std::vector<Camera>& cameras, const std::vector<float>& splatScales,
                    const RenderOptions& options, Sink&& sink, int chunk = 16) {
        if (!splatScales.empty() && splatScales.size() != cameras.size())
            throw std::runtime_error("aniso_amd: one splat scale per camera");
        std::vector<ans_camera> cams(cameras.size());
        for (std::size_t i = 0; i < cameras.size(); ++i) cams[i] = toC(cameras[i]);
        ans_options o{};
        o.sh_degree = options.shDegree;
        o.splat_scale = options.splatScale;
        struct Ctx {
            Sink* sink;
            std::exception_ptr error;
        } ctx{&sink, nullptr};
        auto trampoline = [](void* user, int64_t frame, const uint8_t* rgb, size_t bytes) -> int {
            Ctx* c = static_cast<Ctx*>(user);
            try {
                (*c->sink)(static_cast<std::size_t>(frame), rgb, bytes);
                return 0;
            } catch (...) {
                c->error = std::current_exception();
                return 1;
            }
        };
        const int status = ans_render_path(handle_, static_cast<int64_t>(cams.size()), cams.data(),
                                           splatScales.empty() ? nullptr : splatScales.data(), &o, chunk,
                                           trampoline, &ctx, nullptr);
        if (ctx.error) std::rethrow_exception(ctx.error);
        check(status);
    }

private:
    static ans_camera toC(const Camera& camera) {
        ans_camera c{};
        for (int r = 0; r < 3; ++r)
            for (int k = 0; k < 3; ++k) c.R[3 * r + k] = camera.R[r][k];
        for (int k = 0; k < 3; ++k) c.t[k] = camera.t[k];
        c.fx = camera.fx;
        c.fy = camera.fy;
        c.cx = camera.cx;
        c.cy = camera.cy;
        c.width = camera.width;
        c.height = camera.height;
        return c;
    }
    static void check(int status) {
        if (status != ANS_OK) throw std::runtime_error(std::string("aniso_amd: ") + ans_last_error());
    }
    ans_renderer* handle_ = nullptr;
};

}  // namespace aniso
