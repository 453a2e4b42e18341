// Drives the reference CPU renderer, compiled from its own sources (src/render.cpp, src/sh.cpp,
// src/camera.cpp), through a C entry point the tests can call. TEST INFRASTRUCTURE ONLY: the
// library built from this file is the compiled reference, used to validate the oracle's
// restatement and as the CPU baseline; the product path never loads it.
#include "aniso/render.h"

#include <cstdint>
#include <cstring>

extern "C" {

struct ref_camera {
    double R[9];
    double t[3];
    double fx, fy, cx, cy;
    int width, height;
};

struct ref_stats {
    int64_t visible;
    int64_t pairs;
    double project_ms, sort_ms, blend_ms;
};

// Builds an aniso::Scene from activated arrays (rot is w, x, y, z; sh is [n][coeffs][3]) and
// renders it with aniso::render (render.h:35-36). rgb receives width * height * 3 bytes.
int ref_render(int64_t n, int scene_sh_degree, const float* pos, const float* scale, const float* rot,
               const float* opacity, const float* sh, const ref_camera* cam, int sh_degree, float splat_scale,
               uint8_t* rgb, ref_stats* stats) {
    aniso::Scene scene;
    scene.shDegree = scene_sh_degree;
    scene.gaussians.resize(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; ++i) {
        aniso::Gaussian& g = scene.gaussians[static_cast<size_t>(i)];
        g.position = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
        g.scale = {scale[3 * i], scale[3 * i + 1], scale[3 * i + 2]};
        g.rotation = {rot[4 * i], rot[4 * i + 1], rot[4 * i + 2], rot[4 * i + 3]};
        g.opacity = opacity[i];
    }
    const size_t shCount = static_cast<size_t>(n) * scene.shCoeffsPerChannel() * 3;
    scene.sh.assign(sh, sh + shCount);
    aniso::Camera c;
    c.width = cam->width;
    c.height = cam->height;
    c.fx = cam->fx; c.fy = cam->fy; c.cx = cam->cx; c.cy = cam->cy;
    for (int r = 0; r < 3; ++r) {
        for (int k = 0; k < 3; ++k) c.R[r][k] = cam->R[r * 3 + k];
        c.t[r] = cam->t[r];
    }
    aniso::RenderOptions opt;
    opt.shDegree = sh_degree;
    opt.splatScale = splat_scale;
    aniso::RenderStats st;
    const aniso::Image img = aniso::render(scene, c, &st, opt);
    std::memcpy(rgb, img.rgb.data(), img.rgb.size());
    if (stats) {
        stats->visible = static_cast<int64_t>(st.visible);
        stats->pairs = static_cast<int64_t>(st.tilePairs);
        stats->project_ms = st.projectMs;
        stats->sort_ms = st.sortMs;
        stats->blend_ms = st.blendMs;
    }
    return 0;
}

}  // extern "C"
