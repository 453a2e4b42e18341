// Drives the reference CPU renderer, PLY loader, COLMAP reader and PNG writer, compiled from their
// own sources (src/render.cpp, src/sh.cpp, src/camera.cpp, src/ply_loader.cpp, src/image.cpp),
// through C entry points the
// tests can call. TEST INFRASTRUCTURE ONLY: the
// library built from this file is the compiled reference, used to validate the oracle's
// restatement and as the CPU baseline; the product path never loads it.
#include "aniso/camera.h"
#include "aniso/image.h"
#include "aniso/ply_loader.h"
#include "aniso/render.h"

#include <cstdint>
#include <chrono>
#include <cstring>
#include <exception>

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

// Loads a PLY with aniso::loadPly (ply_loader.h:27). Returns the scene (free with
// ref_scene_free) or NULL with the loader's message in err[errCap]; *ms = the load's wall time.
void* ref_load_ply(const char* path, char* err, size_t errCap, double* ms) {
    try {
        const auto t0 = std::chrono::steady_clock::now();
        auto* s = new aniso::Scene(aniso::loadPly(path));
        if (ms) *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        return s;
    } catch (const std::exception& e) {
        if (err && errCap) {
            std::strncpy(err, e.what(), errCap - 1);
            err[errCap - 1] = 0;
        }
        return nullptr;
    }
}

// Size and arrays of a loaded scene (any array may be NULL): pos/scale [n][3], rot [n][4]
// (w, x, y, z), opacity [n], sh [n][coeffs][3].
void ref_scene_arrays(const void* scene, int64_t* n, int* sh_degree, float* pos, float* scale, float* rot,
                      float* opacity, float* sh) {
    const auto& s = *static_cast<const aniso::Scene*>(scene);
    if (n) *n = static_cast<int64_t>(s.size());
    if (sh_degree) *sh_degree = s.shDegree;
    for (size_t i = 0; i < s.size(); ++i) {
        const aniso::Gaussian& g = s.gaussians[i];
        if (pos) pos[3 * i] = g.position.x, pos[3 * i + 1] = g.position.y, pos[3 * i + 2] = g.position.z;
        if (scale) scale[3 * i] = g.scale.x, scale[3 * i + 1] = g.scale.y, scale[3 * i + 2] = g.scale.z;
        if (rot)
            rot[4 * i] = g.rotation.w, rot[4 * i + 1] = g.rotation.x, rot[4 * i + 2] = g.rotation.y,
            rot[4 * i + 3] = g.rotation.z;
        if (opacity) opacity[i] = g.opacity;
    }
    if (sh) std::memcpy(sh, s.sh.data(), s.sh.size() * sizeof(float));
}

void ref_scene_free(void* scene) { delete static_cast<aniso::Scene*>(scene); }

// aniso::writePng (image.h:22) of rgb[height][width][3]; *ms = its wall time. Returns 0, or -1
// with the message in err[errCap].
int ref_write_png(const char* path, const uint8_t* rgb, int width, int height, char* err, size_t errCap,
                  double* ms) {
    try {
        aniso::Image img(width, height);
        std::memcpy(img.rgb.data(), rgb, img.rgb.size());
        const auto t0 = std::chrono::steady_clock::now();
        aniso::writePng(path, img);
        if (ms) *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        return 0;
    } catch (const std::exception& e) {
        if (err && errCap) {
            std::strncpy(err, e.what(), errCap - 1);
            err[errCap - 1] = 0;
        }
        return -1;
    }
}

// aniso::loadColmapCameras (camera.h:44): *count cameras; if cams is not NULL (room for
// capacity), each as ref_camera plus its NUL-terminated name at names[i * nameCap]. Returns 0,
// or -1 with the message in err[errCap].
int ref_load_colmap(const char* dir, ref_camera* cams, char* names, size_t nameCap, int capacity, int* count,
                    char* err, size_t errCap) {
    try {
        const auto v = aniso::loadColmapCameras(dir);
        *count = static_cast<int>(v.size());
        if (!cams) return 0;
        for (size_t i = 0; i < v.size() && static_cast<int>(i) < capacity; ++i) {
            const aniso::Camera& c = v[i];
            for (int r = 0; r < 3; ++r) {
                for (int k = 0; k < 3; ++k) cams[i].R[3 * r + k] = c.R[r][k];
                cams[i].t[r] = c.t[r];
            }
            cams[i].fx = c.fx, cams[i].fy = c.fy, cams[i].cx = c.cx, cams[i].cy = c.cy;
            cams[i].width = c.width, cams[i].height = c.height;
            std::strncpy(names + i * nameCap, c.imageName.c_str(), nameCap - 1);
            names[i * nameCap + nameCap - 1] = 0;
        }
        return 0;
    } catch (const std::exception& e) {
        if (err && errCap) {
            std::strncpy(err, e.what(), errCap - 1);
            err[errCap - 1] = 0;
        }
        return -1;
    }
}

}  // extern "C"
