// Host side of the rasterizer: device arena, the per-frame launch sequence, and the C ABI
// declared in include/aniso_amd.h.
#include "aniso_amd.h"
#include "engine.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace ans {

thread_local std::string g_lastError;

void hipCheck(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string("HIP ") + what + ": " + hipGetErrorString(e));
}

int bitsFor(uint32_t v) {
    int b = 0;
    while (b < 32 && (1u << b) <= v) ++b;
    return b;
}

CamParams makeCamParams(const ans_camera& c, const ans_options& o, int sceneShDegree) {
    CamParams p{};
    for (int i = 0; i < 9; ++i) p.R[i] = static_cast<float>(c.R[i]);
    for (int i = 0; i < 3; ++i) p.t[i] = static_cast<float>(c.t[i]);
    // Camera centre -R^T t in double, then cast (src/camera.cpp:19-23).
    for (int j = 0; j < 3; ++j)
        p.eye[j] = static_cast<float>(-(c.R[0 * 3 + j] * c.t[0] + c.R[1 * 3 + j] * c.t[1] + c.R[2 * 3 + j] * c.t[2]));
    p.fx = static_cast<float>(c.fx);
    p.fy = static_cast<float>(c.fy);
    p.cx = static_cast<float>(c.cx);
    p.cy = static_cast<float>(c.cy);
    // The Jacobian clamp as src/render.cpp:58-59 derives it.
    p.limX = 1.3f * static_cast<float>(0.5 * c.width / c.fx);
    p.limY = 1.3f * static_cast<float>(0.5 * c.height / c.fy);
    p.width = c.width;
    p.height = c.height;
    p.tileW = o.tile_w;
    p.tileH = o.tile_h;
    p.tileShiftX = bitsFor(static_cast<uint32_t>(o.tile_w)) - 1;
    p.tileShiftY = bitsFor(static_cast<uint32_t>(o.tile_h)) - 1;
    p.tilesX = (c.width + o.tile_w - 1) / o.tile_w;
    p.tilesY = (c.height + o.tile_h - 1) / o.tile_h;
    p.shDegree = std::max(0, std::min(o.sh_degree, sceneShDegree));
    p.shCoeffs = (sceneShDegree + 1) * (sceneShDegree + 1);
    p.splatScale = o.splat_scale;
    return p;
}

ans_options resolveOptions(const ans_options* o) {
    ans_options r{3, 1.0f, 16, 16, 256, 0};
    if (o) {
        r.flags = o->flags;
        r.sh_degree = o->sh_degree;
        r.splat_scale = o->splat_scale;
        if (o->tile_w) r.tile_w = o->tile_w;
        if (o->tile_h) r.tile_h = o->tile_h;
        if (o->batch_depth) r.batch_depth = o->batch_depth;
    }
    const bool tileOk = (r.tile_w == 8 && r.tile_h == 8) || (r.tile_w == 16 && r.tile_h == 16) ||
                        (r.tile_w == 32 && r.tile_h == 16);
    if (!tileOk) throw std::invalid_argument("unsupported tile shape (8x8, 16x16 and 32x16 are supported)");
    if (r.batch_depth != 64 && r.batch_depth != 128 && r.batch_depth != 256)
        throw std::invalid_argument("batch_depth must be 64, 128 or 256");
    if (!(r.splat_scale > 0.0f) || !std::isfinite(r.splat_scale))
        throw std::invalid_argument("splat_scale must be positive and finite");
    return r;
}

void checkCamera(const ans_camera* cam) {
    if (!cam) throw std::invalid_argument("camera is null");
    if (cam->width <= 0 || cam->height <= 0 || cam->width > 16384 || cam->height > 16384)
        throw std::invalid_argument("image size must be in 1..16384");
    if (!(cam->fx > 0) || !(cam->fy > 0)) throw std::invalid_argument("focal lengths must be positive");
}

// ---------------------------------------------------------------------------------------------

FrameState::FrameState() {
    ctl.reserve(1);
    hipCheck(hipMemset(ctl.p, 0, sizeof(FrameCtl)), "hipMemset ctl");
}

void FrameState::reserve(size_t gaussians, size_t tileCount) {
    const size_t m = std::max<size_t>(gaussians, 1);
    rec.reserve(m); tilesTouched.reserve(m); order.reserve(m);
    ranges.reserve(std::max<size_t>(tileCount, 1));
}

void FrameState::reservePairs(size_t cap) {
    if (cap <= pairCap) return;
    tileA.reserve(cap); gaussA.reserve(cap); tileB.reserve(cap); gaussB.reserve(cap);
    pairCap = cap;
}

Engine::Engine() {
    hipCheck(hipGetDevice(&device), "hipGetDevice");
    hipCheck(hipStreamCreateWithFlags(&ownStream, hipStreamNonBlocking), "hipStreamCreate");
    stream = ownStream;
    hipCheck(hipEventCreateWithFlags(&blendDone, hipEventDisableTiming), "hipEventCreate");
    int khz = 0;
    hipCheck(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device), "hipDeviceGetAttribute");
    if (khz <= 0) throw std::runtime_error("device reports no wall clock rate");
    clockKHz = khz;
    hipCheck(hipHostMalloc(reinterpret_cast<void**>(&ctlHost), sizeof(FrameCtl), hipHostMallocDefault),
             "hipHostMalloc");
    hipCheck(hipHostMalloc(reinterpret_cast<void**>(&mirror), sizeof(CountMirror),
                           hipHostMallocMapped | hipHostMallocCoherent),
             "hipHostMalloc");
    std::memset(mirror, 0, sizeof(CountMirror));
    hipCheck(hipHostGetDevicePointer(reinterpret_cast<void**>(&mirrorDev), mirror, 0), "hipHostGetDevicePointer");
    hipCheck(hipHostMalloc(reinterpret_cast<void**>(&stampsHost), sizeof(FrameStamps) * kStampRing,
                           hipHostMallocMapped | hipHostMallocCoherent),
             "hipHostMalloc");
    std::memset(stampsHost, 0, sizeof(FrameStamps) * kStampRing);
    hipCheck(hipHostGetDevicePointer(reinterpret_cast<void**>(&stampsDev), stampsHost, 0), "hipHostGetDevicePointer");
}

Engine::~Engine() {
    // Kernels still in flight write the pinned count mirror and stamp ring: let them finish first
    // (a caller's stream, used by the differentiable path, is drained by ans_diff_destroy).
    if (ownStream) (void)hipStreamSynchronize(ownStream);
    if (ctlHost) (void)hipHostFree(ctlHost);
    if (mirror) (void)hipHostFree(mirror);
    if (stampsHost) (void)hipHostFree(stampsHost);
    if (blendDone) (void)hipEventDestroy(blendDone);
    if (ownStream) (void)hipStreamDestroy(ownStream);
}

void Engine::reserveScratch(size_t n, size_t pairCap) {
    const size_t m = std::max<size_t>(n, 1);
    depthKey.reserve(m); sortKA.reserve(m); sortVA.reserve(m); sortKB.reserve(m); sortVB.reserve(m);
    ttByRank.reserve(m); rectIdx.reserve(m); rectByRank.reserve(m);
    const size_t blocksDepth = (n + kSortBlock * kDepthItems - 1) / (kSortBlock * kDepthItems);
    const size_t blocksDup = (n + kDupTile - 1) / kDupTile;
    const size_t blocksPair = (pairCap + kSortBlock * kTileItems - 1) / (kSortBlock * kTileItems);
    hist.reserve(kRadix * std::max({blocksDepth, blocksPair, size_t(1)}));
    rowTotal.reserve(kRadix);
    blockPairs.reserve(std::max<size_t>(blocksDup, 1));
    blockVisible.reserve(std::max<size_t>(blocksDup, 1));
}

// ---------------------------------------------------------------------------------------------
// The launch sequence of one frame over plain device pointers (FrameBuffers), shared by the
// engine (grow-only buffers) and the entry point that allocates nothing (a fixed scratch array).

void enqueueCountKernels(const FrameBuffers& b, const SceneDev& s, const CamParams& cp, hipStream_t stream,
                         FrameStamps* S, CountMirror* mirror, uint32_t seq) {
    const uint32_t n = static_cast<uint32_t>(s.n);
    const uint32_t tiles = static_cast<uint32_t>(cp.tilesX) * static_cast<uint32_t>(cp.tilesY);
    {
        // The preprocess also clears the frame's control block and tile ranges.
        const uint32_t grid = (std::max<uint32_t>(std::max<uint32_t>(n, tiles), 1) + 255) / 256;
        if (s.raw)
            k_preprocess<false, true><<<grid, 256, 0, stream>>>(static_cast<int>(n), s.pos, s.scale, s.rot, s.opacity,
                                                                s.sh, cp, b.rec, b.tilesTouched, b.depthKey, b.rectIdx, b.ctl,
                                                                b.ranges, tiles, S, s.act);
        else if (s.shPlanar)
            k_preprocess<true, false><<<grid, 256, 0, stream>>>(static_cast<int>(n), s.pos, s.scale, s.rot, s.opacity,
                                                                s.sh, cp, b.rec, b.tilesTouched, b.depthKey, b.rectIdx, b.ctl,
                                                                b.ranges, tiles, S, s.act);
        else
            k_preprocess<false, false><<<grid, 256, 0, stream>>>(static_cast<int>(n), s.pos, s.scale, s.rot,
                                                                 s.opacity, s.sh, cp, b.rec, b.tilesTouched,
                                                                 b.depthKey, b.rectIdx, b.ctl, b.ranges, tiles, S,
                                                                 s.act);
    }
    if (n > 0) {
        // Depth sort: four stable 8-bit passes over all Gaussians (culled ones carry the largest
        // key); ties keep ascending Gaussian index.
        const uint32_t blocks = (n + kSortBlock * kDepthItems - 1) / (kSortBlock * kDepthItems);
        const uint32_t* kin = b.depthKey;
        const uint32_t* vin = nullptr;
        uint32_t* kout[4] = {b.sortKA, b.sortKB, b.sortKA, nullptr};
        uint32_t* vout[4] = {b.sortVA, b.sortVB, b.sortVA, b.order};
        for (int pass = 0; pass < 4; ++pass) {
            k_radixUpsweep<kDepthItems><<<blocks, kSortBlock, 0, stream>>>(kin, false, n, n, 8 * pass, 8, b.hist,
                                                                           blocks, b.ctl, 0, 0);
            k_scanRows<<<kRadix, kSortBlock, 0, stream>>>(b.hist, blocks, blocks, b.rowTotal, b.ctl, false, n,
                                                          kSortBlock * kDepthItems, 0);
            k_radixDownsweep<kDepthItems><<<blocks, kSortBlock, 0, stream>>>(
                kin, vin, kout[pass], vout[pass], false, n, n, 8 * pass, 8, b.hist, b.rowTotal, blocks, b.ctl, 0, 0);
            kin = kout[pass];
            vin = vout[pass];
        }
        const uint32_t blocksDup = (n + kDupTile - 1) / kDupTile;
        k_pairTotals<<<blocksDup, kDupBlock, 0, stream>>>(b.order, b.rectIdx, n, b.ttByRank, b.rectByRank, b.blockPairs,
                                                          b.blockVisible, 0, 0, S, nullptr);
        k_scanPairs<<<1, kScanBlock, 0, stream>>>(b.blockPairs, b.blockVisible, blocksDup,
                                                 static_cast<uint32_t>(b.pairCap), b.ctl, 0, mirror, seq);
    }
}

namespace {

template <int TW, int TH, int BATCH>
void launchBlendFrame(const FrameBuffers& b, const CamParams& cp, const uint32_t* sortedGauss, bool save, bool cull,
                      uint8_t* out8, float* outf, FrameStamps* st, hipStream_t stream) {
    const uint32_t tiles = static_cast<uint32_t>(cp.tilesX) * cp.tilesY;
    if (save)  // frames for the backward pass always cull (culling changes no pixel)
        k_blend<TW, TH, BATCH, true, true><<<tiles, TW * TH, 0, stream>>>(
            b.ranges, sortedGauss, b.rec, cp.width, cp.height, cp.tilesX, out8, outf, b.pixState, b.ctl, st);
    else if (cull)
        k_blend<TW, TH, BATCH, true, false><<<tiles, TW * TH, 0, stream>>>(
            b.ranges, sortedGauss, b.rec, cp.width, cp.height, cp.tilesX, out8, outf, nullptr, b.ctl, st);
    else
        k_blend<TW, TH, BATCH, false, false><<<tiles, TW * TH, 0, stream>>>(
            b.ranges, sortedGauss, b.rec, cp.width, cp.height, cp.tilesX, out8, outf, nullptr, b.ctl, st);
}

template <int TW, int TH>
void launchBlendFrameT(const FrameBuffers& b, const CamParams& cp, const ans_options& o, const uint32_t* sortedGauss,
                       bool save, uint8_t* out8, float* outf, FrameStamps* st, hipStream_t stream) {
    const bool cull = !(o.flags & ANS_FLAG_NO_CULL);
    if (o.batch_depth == 64) launchBlendFrame<TW, TH, 64>(b, cp, sortedGauss, save, cull, out8, outf, st, stream);
    else if (o.batch_depth == 128) launchBlendFrame<TW, TH, 128>(b, cp, sortedGauss, save, cull, out8, outf, st, stream);
    else launchBlendFrame<TW, TH, 256>(b, cp, sortedGauss, save, cull, out8, outf, st, stream);
}

}  // namespace

void enqueueRestKernels(const FrameBuffers& b, uint32_t n, const CamParams& cp, const ans_options& o, bool save,
                        uint8_t* out8, float* outf, hipStream_t stream, FrameStamps* S, hipEvent_t waitBeforeBlend,
                        const uint32_t** sortedTile, const uint32_t** sortedGauss) {
    const uint32_t tiles = static_cast<uint32_t>(cp.tilesX) * static_cast<uint32_t>(cp.tilesY);
    const int tileBits = std::max(1, bitsFor(tiles - 1));
    const int passes = (tileBits + 7) / 8;
    const int digitBits = (tileBits + passes - 1) / passes;
    *sortedTile = b.tileA;
    *sortedGauss = b.gaussA;
    if (n > 0) {
        const uint32_t blocksDup = (n + kDupTile - 1) / kDupTile;
        const uint32_t cap = static_cast<uint32_t>(b.pairCap);
        k_expandPairs<<<blocksDup, kDupBlock, 0, stream>>>(b.order, b.ttByRank, b.rectByRank, n, cp.tilesX,
                                                           b.blockPairs, b.tileA, b.gaussA, b.ctl, 0, 0, 0, nullptr);
        const uint32_t blocksPair =
            static_cast<uint32_t>((b.pairCap + kSortBlock * kTileItems - 1) / (kSortBlock * kTileItems));
        uint32_t* tk[2] = {b.tileA, b.tileB};
        uint32_t* tv[2] = {b.gaussA, b.gaussB};
        for (int pass = 0; pass < passes; ++pass) {
            const int src = pass & 1, dst = src ^ 1;
            const int shift = pass * digitBits;
            const int bits = std::min(digitBits, tileBits - shift);
            k_radixUpsweep<kTileItems><<<blocksPair, kSortBlock, 0, stream>>>(tk[src], true, 0, cap, shift, bits,
                                                                              b.hist, blocksPair, b.ctl, 0, 0);
            k_scanRows<<<1u << bits, kSortBlock, 0, stream>>>(b.hist, blocksPair, blocksPair, b.rowTotal, b.ctl, true,
                                                              cap, kSortBlock * kTileItems, 0);
            k_radixDownsweep<kTileItems><<<blocksPair, kSortBlock, 0, stream>>>(
                tk[src], tv[src], tk[dst], tv[dst], true, 0, cap, shift, bits, b.hist, b.rowTotal, blocksPair, b.ctl,
                0, 0);
        }
        *sortedTile = tk[passes & 1];
        *sortedGauss = tv[passes & 1];
        k_tileRanges<<<static_cast<uint32_t>((b.pairCap + 1023) / 1024), 256, 0, stream>>>(*sortedTile, b.ctl,
                                                                                            b.ranges, 0, 0);
    }
    if (waitBeforeBlend) hipCheck(hipStreamWaitEvent(stream, waitBeforeBlend, 0), "hipStreamWaitEvent");
    if (o.tile_w == 8) launchBlendFrameT<8, 8>(b, cp, o, *sortedGauss, save, out8, outf, S, stream);
    else if (o.tile_w == 32) launchBlendFrameT<32, 16>(b, cp, o, *sortedGauss, save, out8, outf, S, stream);
    else launchBlendFrameT<16, 16>(b, cp, o, *sortedGauss, save, out8, outf, S, stream);
    hipCheck(hipGetLastError(), "kernel launch");
}

// Enqueues one frame on the engine's stream into f. Outputs (either may be null) are device
// pointers.
void Engine::enqueueFrame(FrameState& f, const SceneDev& s, const CamParams& cp, const ans_options& o,
                          uint8_t* out8, float* outf) {
    enqueueCount(f, s, cp, o);
    enqueueRest(f, out8, outf);
}

FrameBuffers Engine::buffers(FrameState& f) {
    FrameBuffers b{};
    b.ctl = f.ctl.p;
    b.rec = f.rec.p;
    b.tilesTouched = f.tilesTouched.p;
    b.order = f.order.p;
    b.tileA = f.tileA.p;
    b.gaussA = f.gaussA.p;
    b.tileB = f.tileB.p;
    b.gaussB = f.gaussB.p;
    b.pairCap = f.pairCap;
    b.ranges = f.ranges.p;
    b.pixState = f.pixState.p;
    b.depthKey = depthKey.p;
    b.sortKA = sortKA.p;
    b.sortVA = sortVA.p;
    b.sortKB = sortKB.p;
    b.sortVB = sortVB.p;
    b.ttByRank = ttByRank.p;
    b.rectIdx = rectIdx.p;
    b.rectByRank = rectByRank.p;
    b.hist = hist.p;
    b.rowTotal = rowTotal.p;
    b.blockPairs = blockPairs.p;
    b.blockVisible = blockVisible.p;
    return b;
}

// First part of a frame: preprocess, depth sort, pair counts and their scan (the pair total is
// in the control block when it completes).
void Engine::enqueueCount(FrameState& f, const SceneDev& s, const CamParams& cp, const ans_options& o) {
    const uint32_t n = static_cast<uint32_t>(s.n);
    const uint32_t tiles = static_cast<uint32_t>(cp.tilesX) * static_cast<uint32_t>(cp.tilesY);
    if (f.pairCap == 0) f.reservePairs(std::max<size_t>(static_cast<size_t>(n) * 8, size_t(1) << 20));
    f.reserve(n, tiles);
    reserveScratch(n, f.pairCap);
    f.tiles = tiles;
    f.n = n;
    f.cp = cp;
    f.opts = o;
    FrameStamps* S = stampsDev + stampNext;
    curStamps = S;
    stampNoPairs[stampNext] = n == 0;  // no pair-counting kernel: the sort stage is empty
    stampNext = (stampNext + 1) % kStampRing;
    stampPending = std::min(stampPending + 1, kStampRing);
    enqueueCountKernels(buffers(f), s, cp, stream, S, mirrorDev, ++countSeq);
}

// Second part: pairs in depth order, a stable sort by tile id, tile ranges, blend. The pair
// buffers may have grown since enqueueCount (their contents are first written here).
void Engine::enqueueRest(FrameState& f, uint8_t* out8, float* outf) {
    reserveScratch(f.n, f.pairCap);
    if (f.savePixels) f.pixState.reserve(static_cast<size_t>(f.cp.width) * f.cp.height);
    enqueueRestKernels(buffers(f), static_cast<uint32_t>(f.n), f.cp, f.opts, f.savePixels, out8, outf, stream,
                       curStamps, waitBeforeBlend, &f.sortedTile, &f.sortedGauss);
    if (recordBlendDone) hipCheck(hipEventRecord(blendDone, stream), "hipEventRecord");
}

bool Engine::readBack(FrameState& f) {
    hipCheck(hipStreamSynchronize(stream), "hipStreamSynchronize");
    // The pair-count scan of each frame published its counts to the mirror (an empty scene has
    // no scan and no pairs).
    f.visible = f.n ? mirror->visible : 0;
    f.pairs = f.n ? mirror->pairs : 0;
    if (mirror->sticky) {
        // Grow to the largest count of the frames that overflowed, which need not be the last
        // frame's.
        const uint32_t pairs = mirror->stickyPairs;
        mirror->sticky = 0;
        mirror->stickyPairs = 0;
        if (pairs >= (1u << 30) - 1) throw std::runtime_error("pair count exceeds 2^30");
        f.reservePairs(static_cast<size_t>(pairs) + pairs / 4 + 1024);
        return true;
    }
    return false;
}

CountMirror Engine::waitCount(const FrameState& f) {
    CountMirror m{};
    if (f.n == 0) return m;  // no pair-count scan was launched: no pairs
    const uint32_t want = countSeq;
    for (uint32_t spin = 1;; ++spin) {
        if (__atomic_load_n(&mirror->seq, __ATOMIC_ACQUIRE) == want) {
            m.pairs = mirror->pairs;
            m.visible = mirror->visible;
            m.overflow = mirror->overflow;
            return m;
        }
        if ((spin & 1023) == 0) {
            // Every so often: has the stream drained without the value arriving? Then read the
            // control block the scan also wrote.
            const hipError_t q = hipStreamQuery(stream);
            if (q == hipSuccess && __atomic_load_n(&mirror->seq, __ATOMIC_ACQUIRE) != want) {
                hipCheck(hipMemcpy(ctlHost, f.ctl.p, sizeof(FrameCtl), hipMemcpyDeviceToHost), "hipMemcpy ctl");
                m.pairs = ctlHost->pairs;
                m.visible = ctlHost->visible;
                m.overflow = ctlHost->overflow;
                return m;
            }
            if (q != hipSuccess && q != hipErrorNotReady) hipCheck(q, "hipStreamQuery");
        }
        __builtin_ia32_pause();
    }
}

// Renders a frame, waiting only for its pair count (the reference reads the same total back,
// gpu_render.cu:324-329). The rest of the frame is enqueued before the wait, with the pair
// buffers as they are, so the GPU does not idle while the host waits: if the pairs exceeded the
// buffers, those kernels skip their work (the control block's overflow flag), the buffers grow
// and the rest of the frame is enqueued again.
int Engine::renderCounted(FrameState& f, const SceneDev& s, const CamParams& cp, const ans_options& o,
                          uint8_t* out8, float* outf) {
    stampPending = 0;  // stage times of this frame
    enqueueCount(f, s, cp, o);
    enqueueRest(f, out8, outf);
    const CountMirror m = waitCount(f);
    f.visible = m.visible;
    f.pairs = m.pairs;
    f.retries = 0;
    if (m.overflow) {
        if (m.pairs >= (1u << 30) - 1) throw std::runtime_error("pair count exceeds 2^30");
        hipCheck(hipStreamSynchronize(stream), "hipStreamSynchronize");  // the skipped kernels have ended
        // This frame's overflow is handled here; overflows of earlier asynchronous frames on the
        // stream stay for their wait to report.
        if (mirror->sticky) mirror->sticky -= 1;
        if (!mirror->sticky) mirror->stickyPairs = 0;
        f.reservePairs(static_cast<size_t>(m.pairs) + m.pairs / 4 + 1024);
        // The flags only stopped the kernels of this frame; with room now, clear them (and the
        // blend's workgroup counter, which the skipped blend advanced).
        hipCheck(hipMemsetAsync(&f.ctl.p->overflow, 0, 12, stream), "hipMemsetAsync");  // overflow, sticky, finished
        f.retries = 1;
        enqueueRest(f, out8, outf);
    }
    return f.retries;
}

void Engine::stageSums(double& a, double& b, double& c, int& frames) {
    // Every frame enqueued since the last readout (the newest kStampRing of them) has completed.
    const int count = std::min(stampPending, kStampRing);
    const double msPerTick = 1.0 / clockKHz;
    for (int k = 0; k < count; ++k) {
        const int slot = (stampNext - 1 - k + kStampRing) % kStampRing;
        const uint64_t* t = stampsHost[slot].t;
        const uint64_t t1 = stampNoPairs[slot] ? t[2] : t[1];
        a += static_cast<double>(t1 - t[0]) * msPerTick;
        b += static_cast<double>(t[2] - t1) * msPerTick;
        c += static_cast<double>(t[3] - t[2]) * msPerTick;
    }
    frames += count;
    stampPending = 0;
}

void Engine::fillStats(const FrameState& f, ans_stats* st, bool synced) {
    if (!st) {
        stampPending = 0;
        return;
    }
    if (!synced) hipCheck(hipStreamSynchronize(stream), "hipStreamSynchronize");
    double a = 0, b = 0, c = 0;
    int frames = 0;
    stageSums(a, b, c, frames);
    frames = std::max(frames, 1);
    st->visible = f.visible;
    st->tile_pairs = f.pairs;
    st->project_ms = a / frames;
    st->sort_ms = b / frames;
    st->blend_ms = c / frames;
    st->pair_capacity = static_cast<int64_t>(f.pairCap);
    st->retries = f.retries;
    st->reserved = 0;
}

// Intermediates of a rendered frame, copied to host buffers (see ans_frame_dump).
void dumpFrame(hipStream_t stream, const FrameState& f, uint32_t* splat_bits, int32_t* rect, uint64_t* keys,
               uint32_t* values, uint32_t* ranges) {
    const size_t n = f.n, P = f.pairs, T = f.tiles;
    hipCheck(hipStreamSynchronize(stream), "hipStreamSynchronize");
    std::vector<SplatRec> rec(n);
    std::vector<float4> r0(n), r1(n), r2(n);
    std::vector<uint2> rc(n);
    std::vector<uint32_t> tt(n);
    auto down = [](void* dst, const void* src, size_t bytes) {
        if (bytes) hipCheck(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hipMemcpy D2H");
    };
    down(tt.data(), f.tilesTouched.p, 4 * n);
    if (splat_bits || keys || rect) {
        down(rec.data(), f.rec.p, sizeof(SplatRec) * n);
        for (size_t i = 0; i < n; ++i) {
            r0[i] = rec[i].geo;
            r1[i] = rec[i].geo2;
            r2[i] = rec[i].col;
            rc[i] = make_uint2(rec[i].meta.x, rec[i].meta.y);
        }
    }
    for (size_t i = 0; i < n; ++i) {
        const bool vis = tt[i] > 0;
        if (splat_bits) {
            const float v[10] = {r0[i].x, r0[i].y, r0[i].z, r0[i].w, r1[i].x, r1[i].y, r2[i].w, r2[i].x, r2[i].y,
                                 r2[i].z};
            if (vis) std::memcpy(splat_bits + 10 * i, v, sizeof v);
            else std::memset(splat_bits + 10 * i, 0, sizeof v);
        }
        if (rect) {
            if (vis) {
                rect[4 * i + 0] = static_cast<int32_t>(rc[i].x & 0xffffu);
                rect[4 * i + 1] = static_cast<int32_t>(rc[i].x >> 16);
                rect[4 * i + 2] = static_cast<int32_t>(rc[i].y & 0xffffu);
                rect[4 * i + 3] = static_cast<int32_t>(rc[i].y >> 16);
            } else {
                rect[4 * i + 0] = 0; rect[4 * i + 1] = 0; rect[4 * i + 2] = -1; rect[4 * i + 3] = -1;
            }
        }
    }
    std::vector<uint32_t> st(P), sg(P);
    if (P && (keys || values)) {
        down(st.data(), f.sortedTile, 4 * P);
        down(sg.data(), f.sortedGauss, 4 * P);
    }
    for (size_t p = 0; p < P; ++p) {
        if (values) values[p] = sg[p];
        if (keys) {
            uint32_t bits;
            std::memcpy(&bits, &r2[sg[p]].w, 4);
            keys[p] = (static_cast<uint64_t>(st[p]) << 32) | bits;
        }
    }
    if (ranges) down(ranges, f.ranges.p, 8 * T);
}

}  // namespace ans

// ---------------------------------------------------------------------------------------------
// C ABI

using namespace ans;

namespace {

void upload(void* dst, const void* src, size_t bytes) {
    if (bytes) hipCheck(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
}

}  // namespace

extern "C" {

int ans_available(void) {
    int count = 0;
    return hipGetDeviceCount(&count) == hipSuccess && count > 0 ? 1 : 0;
}

const char* ans_last_error(void) { return g_lastError.c_str(); }

void ans_activate(int64_t n, const float* log_scale_xyz, const float* opacity_logit, const float* quat_wxyz,
                  float* scale_xyz, float* opacity, float* rot_wxyz) {
    // src/ply_loader.cpp:124-139, in float with std::exp / std::sqrt.
    for (int64_t i = 0; i < n; ++i) {
        for (int j = 0; j < 3; ++j) scale_xyz[3 * i + j] = std::exp(log_scale_xyz[3 * i + j]);
        opacity[i] = 1.0f / (1.0f + std::exp(-opacity_logit[i]));
        const float* q = quat_wxyz + 4 * i;
        const float len = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
        for (int j = 0; j < 4; ++j) rot_wxyz[4 * i + j] = len > 0 ? q[j] / len : (j == 0 ? 1.0f : 0.0f);
    }
}

int ans_renderer_create(int64_t n, int32_t sh_degree, const float* pos_xyz, const float* scale_xyz,
                        const float* rot_wxyz, const float* opacity, const float* sh, ans_renderer** out) {
    return guarded([&] {
        if (!out) throw std::invalid_argument("out is null");
        *out = nullptr;
        if (n < 0 || n > (int64_t(1) << 30)) throw std::invalid_argument("Gaussian count must be in 0..2^30");
        if (sh_degree < 0 || sh_degree > 3) throw std::invalid_argument("sh_degree must be in 0..3");
        if (n > 0 && (!pos_xyz || !scale_xyz || !rot_wxyz || !opacity || !sh))
            throw std::invalid_argument("scene arrays must not be null");
        if (!ans_available()) throw std::runtime_error("no HIP device");
        auto r = std::make_unique<ans_renderer>();
        const size_t m = static_cast<size_t>(std::max<int64_t>(n, 1));
        const size_t coeffs = static_cast<size_t>((sh_degree + 1) * (sh_degree + 1));
        r->pos.reserve(3 * m); r->scale.reserve(3 * m); r->rot.reserve(m); r->opacity.reserve(m);
        r->sh.reserve(coeffs * 3 * m);
        upload(r->pos.p, pos_xyz, sizeof(float) * 3 * n);
        upload(r->scale.p, scale_xyz, sizeof(float) * 3 * n);
        upload(r->rot.p, rot_wxyz, sizeof(float) * 4 * n);
        upload(r->opacity.p, opacity, sizeof(float) * n);
        {
            // Planar SH on the device, [coefficient * 3 + channel][n], so the preprocess reads
            // each coefficient with one coalesced load per wavefront.
            std::vector<float> planar(coeffs * 3 * static_cast<size_t>(n));
            for (int64_t i = 0; i < n; ++i)
                for (size_t j = 0; j < coeffs * 3; ++j) planar[j * n + i] = sh[static_cast<size_t>(i) * coeffs * 3 + j];
            upload(r->sh.p, planar.data(), sizeof(float) * planar.size());
        }
        r->scene = SceneDev{n, sh_degree, r->pos.p, r->scale.p, r->rot.p, r->opacity.p, r->sh.p, true, false};
        r->engine = std::make_unique<Engine>();
        *out = r.release();
        return ANS_OK;
    });
}

int ans_renderer_create_raw(int64_t n, int32_t sh_degree, const float* pos_xyz, const float* log_scale_xyz,
                            const float* quat_wxyz, const float* opacity_logit, const float* sh,
                            ans_renderer** out) {
    return guarded([&] {
        if (n < 0 || n > (int64_t(1) << 30)) throw std::invalid_argument("Gaussian count must be in 0..2^30");
        if (n > 0 && (!log_scale_xyz || !quat_wxyz || !opacity_logit))
            throw std::invalid_argument("scene arrays must not be null");
        std::vector<float> scale(3 * n), opacity(n), rot(4 * n);
        ans_activate(n, log_scale_xyz, opacity_logit, quat_wxyz, scale.data(), opacity.data(), rot.data());
        const int rc = ans_renderer_create(n, sh_degree, pos_xyz, scale.data(), rot.data(), opacity.data(), sh, out);
        if (rc != ANS_OK) throw std::runtime_error(g_lastError);
        return ANS_OK;
    });
}

void ans_renderer_destroy(ans_renderer* r) { delete r; }

int ans_render_float(ans_renderer* r, const ans_camera* cam, const ans_options* options, uint8_t* rgb_out,
                     float* rgbf_out, size_t pixels_capacity, ans_stats* stats) {
    return guarded([&] {
        if (!r) throw std::invalid_argument("renderer is null");
        checkCamera(cam);
        const ans_options o = resolveOptions(options);
        const size_t pixels = static_cast<size_t>(cam->width) * cam->height;
        if ((rgb_out || rgbf_out) && pixels_capacity < pixels)
            throw std::invalid_argument("output buffer smaller than width * height pixels");
        const CamParams cp = makeCamParams(*cam, o, r->scene.shDegree);
        if (cp.tilesX > 65535 || cp.tilesY > 65535) throw std::invalid_argument("too many tiles");
        r->image8.reserve(3 * pixels);
        if (rgbf_out) r->imagef.reserve(3 * pixels);
        Engine& e = *r->engine;
        e.renderCounted(e.frame, r->scene, cp, o, r->image8.p, rgbf_out ? r->imagef.p : nullptr);
        r->last = &e;
        if (rgb_out)
            hipCheck(hipMemcpyAsync(rgb_out, r->image8.p, 3 * pixels, hipMemcpyDeviceToHost, e.stream),
                     "hipMemcpyAsync image");
        if (rgbf_out)
            hipCheck(hipMemcpyAsync(rgbf_out, r->imagef.p, 3 * pixels * sizeof(float), hipMemcpyDeviceToHost,
                                    e.stream),
                     "hipMemcpyAsync image");
        hipCheck(hipStreamSynchronize(e.stream), "hipStreamSynchronize");
        e.fillStats(e.frame, stats, true);
        return ANS_OK;
    });
}

int ans_render(ans_renderer* r, const ans_camera* cam, const ans_options* options, uint8_t* rgb_out,
               size_t rgb_capacity, ans_stats* stats) {
    if (cam && rgb_out && rgb_capacity < static_cast<size_t>(std::max(cam->width, 0)) * std::max(cam->height, 0) * 3) {
        g_lastError = "rgb_out smaller than width * height * 3 bytes";
        return ANS_ERR_ARGUMENT;
    }
    return ans_render_float(r, cam, options, rgb_out, nullptr, rgb_capacity / 3, stats);
}

int ans_render_async(ans_renderer* r, const ans_camera* cam, const ans_options* options, uint8_t* dev_rgb,
                     float* dev_rgbf) {
    return guarded([&] {
        if (!r) throw std::invalid_argument("renderer is null");
        checkCamera(cam);
        const ans_options o = resolveOptions(options);
        const CamParams cp = makeCamParams(*cam, o, r->scene.shDegree);
        if (cp.tilesX > 65535 || cp.tilesY > 65535) throw std::invalid_argument("too many tiles");
        // Consecutive frames alternate between two engines (streams and frame buffers), so one
        // frame's preprocess and sorts run while the previous frame blends; each blend waits for
        // the previous frame's, so frames finish, and write their outputs, in enqueue order.
        Engine* lane = r->asyncNext == 0 ? r->engine.get() : r->lanes[r->asyncNext].get();
        if (!lane) {
            r->lanes[r->asyncNext] = std::make_unique<Engine>();
            lane = r->lanes[r->asyncNext].get();
            lane->frame.reservePairs(r->engine->frame.pairCap);  // as grown so far
        }
        Engine& e = *lane;
        Engine* prev = r->lastAsync;
        e.waitBeforeBlend = prev && prev != &e ? prev->blendDone : nullptr;
        e.recordBlendDone = true;
        try {
            e.enqueueFrame(e.frame, r->scene, cp, o, dev_rgb, dev_rgbf);
        } catch (...) {
            e.waitBeforeBlend = nullptr;
            e.recordBlendDone = false;
            throw;
        }
        e.waitBeforeBlend = nullptr;
        e.recordBlendDone = false;
        r->asyncNext = (r->asyncNext + 1) % ans_renderer::kLanes;
        r->lastAsync = &e;
        r->last = &e;
        return ANS_OK;
    });
}

int ans_render_wait(ans_renderer* r, ans_stats* stats) {
    return guarded([&] {
        if (!r) throw std::invalid_argument("renderer is null");
        bool overflow = false;
        double a = 0, b = 0, c = 0;
        int frames = 0;
        for (int k = 0; k < ans_renderer::kLanes; ++k) {
            Engine* e = k == 0 ? r->engine.get() : r->lanes[k].get();
            if (!e) continue;
            overflow |= e->readBack(e->frame);
            e->frame.retries = 0;
            e->stageSums(a, b, c, frames);
        }
        r->lastAsync = nullptr;
        if (overflow) {
            // Any engine may render the frames again: all get the largest capacity.
            size_t cap = r->engine->frame.pairCap;
            for (int k = 1; k < ans_renderer::kLanes; ++k)
                if (r->lanes[k]) cap = std::max(cap, r->lanes[k]->frame.pairCap);
            r->engine->frame.reservePairs(cap);
            for (int k = 1; k < ans_renderer::kLanes; ++k)
                if (r->lanes[k]) r->lanes[k]->frame.reservePairs(cap);
        }
        if (stats) {
            const Engine& e = r->last ? *r->last : *r->engine;
            frames = std::max(frames, 1);
            stats->visible = e.frame.visible;
            stats->tile_pairs = e.frame.pairs;
            stats->project_ms = a / frames;
            stats->sort_ms = b / frames;
            stats->blend_ms = c / frames;
            stats->pair_capacity = static_cast<int64_t>(e.frame.pairCap);
            stats->retries = 0;
            stats->reserved = 0;
        }
        if (overflow) {
            g_lastError = "pair count exceeded the pair buffers; buffers grown, render again";
            return ANS_ERR_CAPACITY;
        }
        return ANS_OK;
    });
}

int ans_render_autotune(ans_renderer* r, const ans_camera* cam, const ans_options* base, uint8_t* dev_rgb,
                        int32_t repeats, ans_options* best, double* best_ms) {
    return guarded([&] {
        if (!r) throw std::invalid_argument("renderer is null");
        if (!best) throw std::invalid_argument("best is null");
        if (!dev_rgb) throw std::invalid_argument("dev_rgb is null");
        checkCamera(cam);
        const ans_options b = resolveOptions(base);
        const int reps = std::max(1, repeats);
        static const int kTiles[3][2] = {{8, 8}, {16, 16}, {32, 16}};
        static const int kDepths[3] = {64, 128, 256};
        // One frame rendered and waited for: its device time is the sum of its stamp-timed
        // stages. A frame whose pairs overflowed is rendered again (the wait grew the buffers).
        auto frameMs = [&](const ans_options& o) {
            ans_stats st{};
            for (int attempt = 0; attempt < 3; ++attempt) {
                int rc = ans_render_async(r, cam, &o, dev_rgb, nullptr);
                if (rc != ANS_OK) throw std::runtime_error(g_lastError);
                rc = ans_render_wait(r, &st);
                if (rc == ANS_OK) return st.project_ms + st.sort_ms + st.blend_ms;
                if (rc != ANS_ERR_CAPACITY) throw std::runtime_error(g_lastError);
            }
            throw std::runtime_error("pair buffers kept overflowing during autotuning");
        };
        double bestMs = 1e30;
        ans_options bestOpt = b;
        for (const auto& t : kTiles) {
            for (int d : kDepths) {
                ans_options o = b;
                o.tile_w = t[0];
                o.tile_h = t[1];
                o.batch_depth = d;
                for (int k = 0; k < ans_renderer::kLanes; ++k) (void)frameMs(o);  // sizes every lane
                double ms = 1e30;
                for (int k = 0; k < reps; ++k) ms = std::min(ms, frameMs(o));
                if (ms < bestMs) {
                    bestMs = ms;
                    bestOpt = o;
                }
            }
        }
        g_lastError.clear();
        *best = bestOpt;
        if (best_ms) *best_ms = bestMs;
        return ANS_OK;
    });
}

int ans_frame_counts(ans_renderer* r, int64_t* n, int64_t* visible, int64_t* pairs, int64_t* tiles) {
    return guarded([&] {
        if (!r) throw std::invalid_argument("renderer is null");
        const FrameState& f = (r->last ? *r->last : *r->engine).frame;
        if (n) *n = static_cast<int64_t>(f.n);
        if (visible) *visible = f.visible;
        if (pairs) *pairs = f.pairs;
        if (tiles) *tiles = static_cast<int64_t>(f.tiles);
        return ANS_OK;
    });
}

int ans_frame_dump(ans_renderer* r, uint32_t* splat_bits, int32_t* rect, uint64_t* keys, uint32_t* values,
                   uint32_t* ranges) {
    return guarded([&] {
        if (!r) throw std::invalid_argument("renderer is null");
        const Engine& e = r->last ? *r->last : *r->engine;
        dumpFrame(e.stream, e.frame, splat_bits, rect, keys, values, ranges);
        return ANS_OK;
    });
}

}  // extern "C"
