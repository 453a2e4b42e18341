// Device arena and per-frame launch sequence of the rasterizer (host side).
#pragma once

#include "aniso_amd.h"
#include "forward.hpp"

#include <cstddef>

#ifndef ANS_ASYNC_LANES
#define ANS_ASYNC_LANES 3
#endif
#include <memory>
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

// A scene in device memory. Activated (render API; layout of gpu_render.cu:243-248 with planar SH)
// or raw (the differentiable path: log-scales, opacity logits, unnormalized quaternions, SH
// [n][coefficient][channel], as a training loop holds them).
struct SceneDev {
    int64_t n = 0;
    int shDegree = 0;
    const float* pos = nullptr;
    const float* scale = nullptr;    // activated: standard deviations; raw: log-scales
    const float4* rot = nullptr;     // activated: unit (w, x, y, z); raw: unnormalized
    const float* opacity = nullptr;  // activated: opacity; raw: logit
    const float* sh = nullptr;
    bool shPlanar = false;  // [coefficient * 3 + channel][n] instead of [n][coefficient][channel]
    bool raw = false;
    bool act = false;  // raw with activated scales and opacities (ans_params.activated)
};

// Device buffers of one single-view frame as plain pointers: the frame's own (control block,
// splat records, tile counts, depth order, pair double buffers of pairCap slots, tile ranges,
// pixel state if the backward pass will read it) and the sort and scan scratch.
struct FrameBuffers {
    FrameCtl* ctl;
    SplatRec* rec;
    uint32_t *tilesTouched, *order;
    uint32_t *tileA, *gaussA, *tileB, *gaussB;
    size_t pairCap;
    uint2* ranges;
    uint2* pixState;
    uint32_t *depthKey, *sortKA, *sortVA, *sortKB, *sortVB, *ttByRank;
    uint2 *rectIdx, *rectByRank;
    uint32_t *hist, *rowTotal, *blockPairs, *blockVisible;
};

// A frame's launches: preprocess, depth sort and pair count (published to the control block and,
// with mirror, to pinned host memory under number seq); then pair expansion, tile sort, tile
// ranges and the blend (save: with pixel state for the backward pass), after waitBeforeBlend if
// given. Kernel launches only.
void enqueueCountKernels(const FrameBuffers& b, const SceneDev& s, const CamParams& cp, hipStream_t stream,
                         FrameStamps* stamps, CountMirror* mirror, uint32_t seq);
void enqueueRestKernels(const FrameBuffers& b, uint32_t n, const CamParams& cp, const ans_options& o, bool save,
                        uint8_t* out8, float* outf, hipStream_t stream, FrameStamps* stamps,
                        hipEvent_t waitBeforeBlend, const uint32_t** sortedTile, const uint32_t** sortedGauss);

CamParams makeCamParams(const ans_camera& c, const ans_options& o, int sceneShDegree);
ans_options resolveOptions(const ans_options* o);
void checkCamera(const ans_camera* cam);

// Everything one frame produces that its blend and its backward pass read.
struct FrameState {
    DevBuf<FrameCtl> ctl;
    DevBuf<SplatRec> rec;
    DevBuf<uint32_t> tilesTouched;
    DevBuf<uint32_t> order;  // Gaussians by depth (rank -> index)
    DevBuf<uint32_t> tileA, gaussA, tileB, gaussB;
    size_t pairCap = 0;
    DevBuf<uint2> ranges;
    bool savePixels = false;  // the blend records pixState (frames the backward pass will read)
    DevBuf<uint2> pixState;   // per pixel: final transmittance bits, last blended list position + 1

    size_t n = 0, tiles = 0;
    uint32_t visible = 0, pairs = 0;
    int retries = 0;
    CamParams cp{};
    ans_options opts{};
    const uint32_t* sortedTile = nullptr;
    const uint32_t* sortedGauss = nullptr;

    FrameState();
    void reserve(size_t gaussians, size_t tileCount);
    void reservePairs(size_t cap);
};

// Per-stream working set: scratch shared by all frames plus the render API's own frame.
struct Engine {
    int device = 0;
    hipStream_t ownStream = nullptr;
    hipStream_t stream = nullptr;
    // Stage timestamps of the last kStampRing frames, written by the frames' kernels into pinned
    // host memory; stats average the frames enqueued since the previous readout.
    static constexpr int kStampRing = 64;
    FrameStamps* stampsHost = nullptr;  // pinned, coherent
    FrameStamps* stampsDev = nullptr;   // its device address
    bool stampNoPairs[kStampRing] = {};
    int stampNext = 0, stampPending = 0;
    FrameStamps* curStamps = nullptr;  // stamps of the frame being enqueued
    double clockKHz = 0;               // wall clock rate
    FrameCtl* ctlHost = nullptr;       // pinned
    // Ordering of the blends of frames pipelined over two engines (the asynchronous render API):
    // with waitBeforeBlend set, the next frame's blend waits for that event (the other engine's
    // previous blend); with recordBlendDone, blendDone is recorded after the blend.
    hipEvent_t blendDone = nullptr;
    hipEvent_t waitBeforeBlend = nullptr;
    bool recordBlendDone = false;
    CountMirror* mirror = nullptr;     // pinned, coherent: written by the pair-count scan
    CountMirror* mirrorDev = nullptr;  // its device address
    uint32_t countSeq = 0;             // number of the last frame whose count was enqueued

    DevBuf<uint32_t> depthKey, sortKA, sortVA, sortKB, sortVB;
    DevBuf<uint32_t> ttByRank;      // tile counts in depth order
    DevBuf<uint2> rectIdx;          // packed tile rects by Gaussian index
    DevBuf<uint2> rectByRank;       // tile rects in depth order
    DevBuf<uint32_t> hist;          // per-block digit histograms, scanned in place
    DevBuf<uint32_t> rowTotal;      // per-digit totals of one pass
    DevBuf<uint32_t> blockPairs;    // pairs per expansion block, scanned in place
    DevBuf<uint32_t> blockVisible;  // visible Gaussians per expansion block
    DevBuf<float> partial;          // backward: per-pair partial gradients, 12-float records
    DevBuf<float> splatGrad;        // backward: per-Gaussian splat-record gradients, 9 floats
    DevBuf<uint32_t> slotStart;     // backward: first pair slot per Gaussian, index order
    FrameState frame;

    Engine();
    ~Engine();
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;

    void reserveScratch(size_t n, size_t pairCap);
    void enqueueFrame(FrameState& f, const SceneDev& s, const CamParams& cp, const ans_options& o, uint8_t* out8,
                      float* outf);
    void enqueueCount(FrameState& f, const SceneDev& s, const CamParams& cp, const ans_options& o);
    void enqueueRest(FrameState& f, uint8_t* out8, float* outf);
    FrameBuffers buffers(FrameState& f);
    // Waits for the stream; true if the pair buffers overflowed in any frame since the last call
    // (they have then been grown).
    bool readBack(FrameState& f);
    // Waits until the pair count of the frame enqueued last has been published (not for the rest
    // of the frame) and returns it.
    CountMirror waitCount(const FrameState& f);
    int renderCounted(FrameState& f, const SceneDev& s, const CamParams& cp, const ans_options& o, uint8_t* out8,
                      float* outf);
    // Fills st (if not null) for frame f; synced: the caller has waited for the stream, else this
    // waits.
    void fillStats(const FrameState& f, ans_stats* st, bool synced = false);
    // Adds the stage times (ms) of the frames enqueued since the last readout to a, b, c and
    // their number to frames, and starts a new readout.
    void stageSums(double& a, double& b, double& c, int& frames);
    void enqueueBackward(FrameState& f, const SceneDev& s, const float* gradImage, float* gMeans, float* gScales,
                         float* gQuats, float* gOpacity, float* gSh);
};

// Runs fn, turning exceptions into the C ABI's status codes and ans_last_error().
template <class Fn>
int guarded(Fn&& fn) {
    try {
        g_lastError.clear();
        return fn();
    } catch (const std::invalid_argument& e) {
        g_lastError = e.what();
        return ANS_ERR_ARGUMENT;
    } catch (const std::exception& e) {
        g_lastError = e.what();
        return ANS_ERR_HIP;
    } catch (...) {
        g_lastError = "unknown error";
        return ANS_ERR_HIP;
    }
}

struct ViewBatch;  // batched multi-view state (batch.hip)
struct ViewBatchDeleter {
    void operator()(ViewBatch* b) const;
};

}  // namespace ans

// The render handle: the scene in device memory (planar SH), the engine, and the batched
// multi-view state once used.
struct ans_renderer {
    ans::SceneDev scene;
    ans::DevBuf<float> pos, scale, opacity, sh;
    ans::DevBuf<float4> rot;
    ans::DevBuf<uint8_t> image8;
    ans::DevBuf<float> imagef;
    std::unique_ptr<ans::Engine> engine;
    // Asynchronous frames rotate over kLanes engines (engine, then lanes[1..], created on first
    // use), each with its own stream and frame buffers, so that one frame's preprocess and sorts
    // overlap the previous frame's blend; blends still run in enqueue order.
    static constexpr int kLanes = ANS_ASYNC_LANES;
    std::unique_ptr<ans::Engine> lanes[kLanes];
    int asyncNext = 0;
    ans::Engine* lastAsync = nullptr;  // engine of the last asynchronous frame not yet waited for
    ans::Engine* last = nullptr;       // engine of the last frame enqueued (counts, dumps)
    std::unique_ptr<ans::ViewBatch, ans::ViewBatchDeleter> batch;
};
