// Device arena and per-frame launch sequence of the rasterizer (host side).
#pragma once

#include "aniso_amd.h"
#include "forward.hpp"

#include <cstddef>
#include <stdexcept>
#include <string>

namespace ans {

extern thread_local std::string g_lastError;
void hipCheck(hipError_t e, const char* what);
int bitsFor(uint32_t v);

// A device buffer that only grows (the reference's Buffer<T>, gpu_render.cu:23-37).
template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t cap = 0;
    void reserve(size_t n) {
        if (n <= cap) return;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        hipCheck(hipMalloc(reinterpret_cast<void**>(&p), std::max<size_t>(n, 1) * sizeof(T)), "hipMalloc");
        cap = n;
    }
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

// Activated scene in device memory (layout of gpu_render.cu:243-248: pos f3, scale f3,
// rot f4 (w first), opacity f, sh [n][coeffs][3]).
struct SceneDev {
    int64_t n = 0;
    int shDegree = 0;
    const float* pos = nullptr;
    const float* scale = nullptr;
    const float4* rot = nullptr;
    const float* opacity = nullptr;
    const float* sh = nullptr;
    bool shPlanar = false;  // [coefficient * 3 + channel][n] instead of [n][coefficient][channel]
};

CamParams makeCamParams(const ans_camera& c, const ans_options& o, int sceneShDegree);
ans_options resolveOptions(const ans_options* o);
void checkCamera(const ans_camera* cam);

struct LastFrame {
    size_t n = 0, tiles = 0;
    uint32_t visible = 0, pairs = 0;
    int retries = 0;
    CamParams cp{};
    const uint32_t* sortedTile = nullptr;
    const uint32_t* sortedGauss = nullptr;
};

// Per-stream working set: every per-Gaussian, per-pair and per-tile buffer of a frame.
struct Engine {
    int device = 0;
    hipStream_t ownStream = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {};

    DevBuf<FrameCtl> ctl;
    FrameCtl* ctlHost = nullptr;
    DevBuf<float4> rec0, rec1, rec2;
    DevBuf<uint2> rect;
    DevBuf<uint32_t> tilesTouched, depthKey, pairStart;
    DevBuf<uint32_t> sortKA, sortVA, sortKB, sortVB;
    DevBuf<uint32_t> tileA, gaussA, tileB, gaussB;
    size_t pairCap = 0;
    DevBuf<uint2> ranges;
    DevBuf<uint32_t> hist;        // per-block digit histograms, scanned in place
    DevBuf<uint32_t> rowTotal;    // per-digit totals of one pass
    DevBuf<uint32_t> blockPairs;  // pairs per expansion block, scanned in place
    DevBuf<uint32_t> blockVisible;  // visible Gaussians per expansion block
    LastFrame last;

    Engine();
    ~Engine();
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;

    void reserveGaussians(size_t n);
    void reservePairs(size_t cap);
    void reserveScratch(size_t n);
    void enqueueFrame(const SceneDev& s, const CamParams& cp, const ans_options& o, uint8_t* out8, float* outf);
    void launchBlend(const CamParams& cp, const ans_options& o, uint8_t* out8, float* outf);
    template <int TW, int TH>
    void launchBlendT(const CamParams& cp, int batch, bool cull, uint8_t* out8, float* outf);
    template <int TW, int TH, int BATCH>
    void launchBlendB(const CamParams& cp, bool cull, uint8_t* out8, float* outf);
    int renderSync(const SceneDev& s, const CamParams& cp, const ans_options& o, uint8_t* out8, float* outf);
    void fillStats(ans_stats* st);
};

}  // namespace ans
