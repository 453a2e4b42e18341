// Batched differentiable frames: one parameter set rendered from many cameras of uneven sizes
// by one sequence of batched launches (the batched render's, batch.hip, on raw parameters, with
// float images and per-pixel state saved), and its backward: one slot-numbering scan over all
// views, one blend-adjoint launch over every (view, tile) item, and one per-Gaussian launch that
// adds every view's gradient in view order. Both are captured as HIP graphs and replayed.
//
// There is no counterpart in the reference: its training loop renders and differentiates one
// view per step through a third-party rasterizer (pipeline/train_gsplat.py:170-190, the render
// at :96-102). Each view here equals a single-view differentiable forward of that camera, and
// the gradients equal the sum over views, in view order, of the single-view backward gradients.
#include "aniso_amd.h"
#include "diff.hpp"

#include <algorithm>
#include <stdexcept>
#include <string>
#include <vector>

namespace ans {

DiffViewsFrame::DiffViewsFrame() {
    hipCheck(hipEventCreateWithFlags(&staged, hipEventDisableTiming), "hipEventCreate");
}

DiffViewsFrame::~DiffViewsFrame() {
    dropBwdGraphs();
    if (staged) {
        (void)hipEventSynchronize(staged);
        (void)hipEventDestroy(staged);
    }
}

void DiffViewsFrame::dropBwdGraphs() {
    for (auto& kv : bwdGraphs) (void)hipGraphExecDestroy(kv.second);
    bwdGraphs.clear();
}

namespace {

constexpr int kBwdBatch = 128;  // LDS batch of the blend adjoint (the single-view adjoint's)

template <int TW, int TH>
void launchBlendBackwardViews(DiffViewsFrame& F, DiffViewsScratch& S, const TableLayout& T, hipStream_t st) {
    ViewBatch& B = F.B;
    const Layout& L = F.L;
    const uint8_t* tb = B.table.p;
    k_blendBackwardViews<TW, TH, kBwdBatch><<<L.totalTiles, TW * TH, 0, st>>>(
        B.itemOrder.p, L.maxTiles, B.ranges.p, B.sortedGauss, static_cast<uint32_t>(L.pairCap), B.rec.p, L.n,
        S.slotStart.p, B.pixState.p, reinterpret_cast<const uint64_t*>(tb + T.pixOff),
        reinterpret_cast<const CamParams*>(tb + T.cams), reinterpret_cast<const float* const*>(tb + T.grads),
        reinterpret_cast<const uint64_t*>(tb + T.pairBase), S.partial.p);
}

// The backward's launches (capturable): pair slots per view in Gaussian index order, the blend
// adjoint over every (view, tile) item, the per-Gaussian adjoint summing the views.
void enqueueBackwardViews(DiffViewsFrame& F, DiffViewsScratch& S, hipStream_t st) {
    const Layout& L = F.L;
    ViewBatch& B = F.B;
    const TableLayout T(L.views);
    const uint32_t n = L.n, V = static_cast<uint32_t>(L.views);
    const uint32_t blocks = (n + 255) / 256;
    k_slotBlockSums<<<dim3(blocks, V), 256, 0, st>>>(n, B.tilesTouched.p, S.blockSums.p, n, blocks);
    k_scanBlockSums<<<dim3(1, V), kScanBlock, 0, st>>>(S.blockSums.p, blocks, blocks);
    k_slotStart<<<dim3(blocks, V), 256, 0, st>>>(n, B.tilesTouched.p, S.blockSums.p, S.slotStart.p, n, blocks);
    if (L.opts.tile_w == 8) launchBlendBackwardViews<8, 8>(F, S, T, st);
    else if (L.opts.tile_w == 32) launchBlendBackwardViews<32, 16>(F, S, T, st);
    else launchBlendBackwardViews<16, 16>(F, S, T, st);
    k_gaussBackwardViews<<<(n + kWave - 1) / kWave, kWave, 0, st>>>(
        static_cast<int>(n), L.views, reinterpret_cast<const DiffParamsDev*>(B.table.p + T.params),
        reinterpret_cast<const CamParams*>(B.table.p + T.cams), B.tilesTouched.p, S.slotStart.p,
        reinterpret_cast<const uint64_t*>(B.table.p + T.pairBase), S.partial.p);
    hipCheck(hipGetLastError(), "kernel launch");
}

template <class T>
bool growScratch(DevBuf<T>& b, size_t n) {
    if (n <= b.cap) return false;
    b.reserve(n);
    return true;
}

DiffParamsDev paramsDev(const SceneDev& s, const ans_grads* g, float* screenGrads) {
    DiffParamsDev P{};
    P.pos = s.pos;
    P.scale = s.scale;
    P.rot = s.rot;
    P.opacity = s.opacity;
    P.sh = s.sh;
    if (g) {
        P.gMeans = g->means;
        P.gScales = g->log_scales;
        P.gQuats = g->quats;
        P.gOpacity = g->opacity_logits;
        P.gSh = g->sh;
    }
    P.screenGrads = reinterpret_cast<float2*>(screenGrads);
    P.shCoeffs = (s.shDegree + 1) * (s.shDegree + 1);
    P.act = s.act ? 1 : 0;
    return P;
}

template <class Fn>
int guardedViews(Fn&& fn) {
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

DiffViewsFrame& viewFrameOf(ans_diff* d, int64_t id) {
    if (!d) throw std::invalid_argument("handle is null");
    auto it = d->viewFrames.find(id);
    if (it == d->viewFrames.end()) throw std::invalid_argument("unknown or released batched frame");
    return *it->second;
}

}  // namespace

}  // namespace ans

using namespace ans;

extern "C" {

int ans_diff_views_forward(ans_diff* d, const ans_params* params, int32_t n_views, const ans_camera* cams,
                           const ans_options* options, float* const* images, void* stream, int64_t* frame,
                           ans_stats* stats) {
    return guardedViews([&] {
        if (!d) throw std::invalid_argument("handle is null");
        if (n_views < 1 || n_views > kMaxViews) throw std::invalid_argument("n_views must be in 1..256");
        if (!cams || !images) throw std::invalid_argument("cameras and images must not be null");
        const SceneDev s = sceneOfParams(params);
        const ans_options o = resolveOptions(options);
        std::vector<CamParams> cps(n_views);
        for (int v = 0; v < n_views; ++v) {
            checkCamera(&cams[v]);
            if (!images[v]) throw std::invalid_argument("image pointer must not be null");
            cps[v] = makeCamParams(cams[v], o, s.shDegree);
            if (cps[v].tilesX > 65535 || cps[v].tilesY > 65535) throw std::invalid_argument("too many tiles");
        }
        std::unique_ptr<DiffViewsFrame> f;
        if (!d->viewPool.empty()) {
            f = std::move(d->viewPool.back());
            d->viewPool.pop_back();
        } else {
            f = std::make_unique<DiffViewsFrame>();
        }
        ViewBatch& B = f->B;
        const DiffParamsDev P = paramsDev(s, nullptr, nullptr);
        BatchIO io;
        io.outf = images;
        io.params = &P;
        int retries = 0;
        Layout L{};
        while (true) {
            L = makeLayout(B, s, n_views, cps, o, true);
            reserveArena(B, L);
            B.pairCap = L.pairCap;
            submitBatch(B, s, L, cps, io, static_cast<hipStream_t>(stream), 0);
            if (!waitBatchCounts(B, L, cps, 0)) break;
            ++retries;
        }
        B.lastRetries = retries;
        f->L = L;
        f->cps = cps;
        f->n = s.n;
        f->shDegree = s.shDegree;
        f->act = s.act;
        if (stats) {
            hipCheck(hipEventSynchronize(B.ev[0][1]), "hipEventSynchronize");
            hipCheck(hipEventElapsedTime(&B.lastMs, B.ev[0][0], B.ev[0][1]), "hipEventElapsedTime");
            for (int v = 0; v < n_views; ++v) {
                ans_stats& st = stats[v];
                st.visible = B.lastCtl[v].visible;
                st.tile_pairs = B.lastCtl[v].pairs;
                st.project_ms = 0.0;
                st.sort_ms = 0.0;
                st.blend_ms = B.lastMs;
                st.pair_capacity = static_cast<int64_t>(B.pairCap);
                st.retries = retries;
                st.reserved = 0;
            }
        }
        const int64_t id = d->next++;
        d->viewFrames[id] = std::move(f);
        if (frame) *frame = id;
        return ANS_OK;
    });
}

int ans_diff_views_backward(ans_diff* d, int64_t frame, const ans_params* params, const float* const* grad_images,
                            const ans_grads* grads, float* screen_grads, void* stream) {
    return guardedViews([&] {
        DiffViewsFrame& F = viewFrameOf(d, frame);
        const SceneDev s = sceneOfParams(params);
        if (s.n != F.n || s.shDegree != F.shDegree || s.act != F.act)
            throw std::invalid_argument("parameters differ from the forward's (count, SH degree or activation)");
        const int V = F.L.views;
        if (!grad_images || !grads || (s.n > 0 && (!grads->means || !grads->log_scales || !grads->quats ||
                                                   !grads->opacity_logits || !grads->sh)))
            throw std::invalid_argument("gradient pointers must not be null");
        for (int v = 0; v < V; ++v)
            if (!grad_images[v]) throw std::invalid_argument("image gradient pointer must not be null");
        if (reinterpret_cast<uintptr_t>(grads->quats) % 16 != 0)
            throw std::invalid_argument("grads quats must be 16-byte aligned");
        if (reinterpret_cast<uintptr_t>(screen_grads) % 8 != 0)
            throw std::invalid_argument("screen_grads must be 8-byte aligned");
        if (s.n == 0) return ANS_OK;
        hipStream_t st = static_cast<hipStream_t>(stream);
        ViewBatch& B = F.B;
        DiffViewsScratch& S = d->viewScratch;
        // Partial records of view v from pair pairBase[v]: the views' pair counts in order.
        std::vector<uint64_t> base(V);
        uint64_t total = 0;
        for (int v = 0; v < V; ++v) {
            base[v] = total;
            total += B.lastCtl[v].pairs;
        }
        const size_t n = static_cast<size_t>(F.n);
        const size_t blocks = (n + 255) / 256;
        bool moved = false;
        moved |= growScratch(S.slotStart, static_cast<size_t>(V) * n);
        moved |= growScratch(S.blockSums, static_cast<size_t>(V) * blocks);
        moved |= growScratch(S.partial, std::max<uint64_t>(total, 1) * kPartialStride);
        if (moved) ++S.generation;
        // The table's backward part: image gradients, pair bases, parameter and gradient arrays
        // (one contiguous region at its end), staged in the second pinned buffer.
        const TableLayout T(V);
        hipCheck(hipEventSynchronize(F.staged), "hipEventSynchronize");
        uint8_t* th = B.tableHost[1];
        for (int v = 0; v < V; ++v) {
            reinterpret_cast<const float**>(th + T.grads)[v] = grad_images[v];
            reinterpret_cast<uint64_t*>(th + T.pairBase)[v] = base[v];
        }
        *reinterpret_cast<DiffParamsDev*>(th + T.params) = paramsDev(s, grads, screen_grads);
        hipCheck(hipMemcpyAsync(B.table.p + T.grads, th + T.grads, T.bytes - T.grads, hipMemcpyHostToDevice, st),
                 "hipMemcpyAsync table");
        hipCheck(hipEventRecord(F.staged, st), "hipEventRecord");
        if (F.L.opts.flags & ANS_FLAG_NO_GRAPH) {
            enqueueBackwardViews(F, S, st);
            return ANS_OK;
        }
        const DiffViewsFrame::BwdKey key{V, F.n, F.L.opts.tile_w, F.L.opts.batch_depth, F.L.maxTiles,
                                         F.L.totalTiles, F.L.pairCap, B.generation, S.generation, F.L.pixels};
        auto it = F.bwdGraphs.find(key);
        if (it == F.bwdGraphs.end()) {
            hipGraph_t g = nullptr;
            hipCheck(hipStreamBeginCapture(B.stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
            try {
                enqueueBackwardViews(F, S, B.stream);
            } catch (...) {
                (void)hipStreamEndCapture(B.stream, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            hipCheck(hipStreamEndCapture(B.stream, &g), "hipStreamEndCapture");
            hipGraphExec_t exec = nullptr;
            const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            hipCheck(e, "hipGraphInstantiate");
            it = F.bwdGraphs.emplace(key, exec).first;
        }
        hipCheck(hipGraphLaunch(it->second, st), "hipGraphLaunch");
        return ANS_OK;
    });
}

}  // extern "C"

namespace {

// 1 - final transmittance per pixel of one view.
__global__ __launch_bounds__(256) void k_alphaFromState(size_t pixels, const uint2* __restrict__ pixState,
                                                        float* __restrict__ alpha) {
    const size_t p = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    if (p < pixels) alpha[p] = 1.0f - __uint_as_float(pixState[p].x);
}

// Per (view, Gaussian): splat radius in pixels, 0 if culled in that view.
__global__ __launch_bounds__(256) void k_radiiViews(size_t count, const SplatRec* __restrict__ rec,
                                                    const uint32_t* __restrict__ tilesTouched,
                                                    int32_t* __restrict__ radii) {
    const size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    if (i < count) radii[i] = tilesTouched[i] ? static_cast<int32_t>(rec[i].meta.w) : 0;
}

}  // namespace

extern "C" {

int ans_diff_views_alpha(ans_diff* d, int64_t frame, float* const* alphas, void* stream) {
    return guardedViews([&] {
        DiffViewsFrame& F = viewFrameOf(d, frame);
        if (!alphas) throw std::invalid_argument("alphas is null");
        uint64_t off = 0;
        for (int v = 0; v < F.L.views; ++v) {
            const size_t pixels = static_cast<size_t>(F.cps[v].width) * F.cps[v].height;
            if (!alphas[v]) throw std::invalid_argument("alpha pointer must not be null");
            k_alphaFromState<<<static_cast<uint32_t>((pixels + 255) / 256), 256, 0, static_cast<hipStream_t>(stream)>>>(
                pixels, F.B.pixState.p + off, alphas[v]);
            off += pixels;
        }
        hipCheck(hipGetLastError(), "kernel launch");
        return ANS_OK;
    });
}

int ans_diff_views_radii(ans_diff* d, int64_t frame, int32_t* radii, void* stream) {
    return guardedViews([&] {
        DiffViewsFrame& F = viewFrameOf(d, frame);
        const size_t count = static_cast<size_t>(F.L.views) * static_cast<size_t>(F.n);
        if (!radii && count) throw std::invalid_argument("radii is null");
        if (count) {
            k_radiiViews<<<static_cast<uint32_t>((count + 255) / 256), 256, 0, static_cast<hipStream_t>(stream)>>>(
                count, F.B.rec.p, F.B.tilesTouched.p, radii);
            hipCheck(hipGetLastError(), "kernel launch");
        }
        return ANS_OK;
    });
}

int ans_diff_views_counts(ans_diff* d, int64_t frame, int32_t view, int64_t* visible, int64_t* pairs,
                          int64_t* tiles) {
    return guardedViews([&] {
        DiffViewsFrame& F = viewFrameOf(d, frame);
        if (view < 0 || view >= F.L.views) throw std::invalid_argument("view out of range");
        if (visible) *visible = F.B.lastCtl[view].visible;
        if (pairs) *pairs = F.B.lastCtl[view].pairs;
        if (tiles) *tiles = static_cast<int64_t>(F.cps[view].tilesX) * F.cps[view].tilesY;
        return ANS_OK;
    });
}

int ans_diff_views_dump(ans_diff* d, int64_t frame, int32_t view, uint32_t* values, uint32_t* ranges) {
    return guardedViews([&] {
        DiffViewsFrame& F = viewFrameOf(d, frame);
        if (view < 0 || view >= F.L.views) throw std::invalid_argument("view out of range");
        const ViewBatch& B = F.B;
        const size_t P = B.lastCtl[view].pairs;
        const size_t T = static_cast<size_t>(F.cps[view].tilesX) * F.cps[view].tilesY;
        hipCheck(hipDeviceSynchronize(), "hipDeviceSynchronize");
        if (values && P)
            hipCheck(hipMemcpy(values, B.sortedGauss + static_cast<size_t>(view) * B.pairCap, 4 * P,
                               hipMemcpyDeviceToHost),
                     "hipMemcpy values");
        if (ranges && T)
            hipCheck(hipMemcpy(ranges, B.ranges.p + static_cast<size_t>(view) * B.maxTiles, 8 * T,
                               hipMemcpyDeviceToHost),
                     "hipMemcpy ranges");
        return ANS_OK;
    });
}

}  // extern "C"
