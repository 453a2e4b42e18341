// Differentiable path: forward frames kept for their backward pass, and the backward launch
// sequence (blend adjoint, then per-Gaussian reduction and projection adjoint). C ABI in
// include/aniso_amd.h.
#include "aniso_amd.h"
#include "diff.hpp"

#include <map>
#include <memory>
#include <vector>

namespace ans {

void dumpFrame(hipStream_t stream, const FrameState& f, uint32_t* splat_bits, int32_t* rect, uint64_t* keys,
               uint32_t* values, uint32_t* ranges);

template <int TW, int TH>
static void launchBlendBackward(Engine& e, FrameState& f, const float* gradImage) {
    constexpr int kBatch = 128;
    const CamParams& cp = f.cp;
    const uint32_t tiles = static_cast<uint32_t>(cp.tilesX) * cp.tilesY;
    k_blendBackward<TW, TH, kBatch><<<tiles, TW * TH, 0, e.stream>>>(
        f.ranges.p, f.sortedGauss, f.rec.p, e.slotStart.p, f.pixState.p, cp.width, cp.height, cp.tilesX, gradImage,
        e.partial.p, f.ctl.p);
}

// Per Gaussian: its splat radius in pixels (0 if culled).
static __global__ __launch_bounds__(256) void k_radii(uint32_t n, const SplatRec* __restrict__ rec,
                                                      const uint32_t* __restrict__ tilesTouched,
                                                      int32_t* __restrict__ radii) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) radii[i] = tilesTouched[i] ? static_cast<int32_t>(rec[i].meta.w) : 0;
}

// Per Gaussian: dL/du, dL/dv (columns 0 and 1 of the splat-record gradient).
static __global__ __launch_bounds__(256) void k_screenGrads(uint32_t n, const float* __restrict__ splatGrad,
                                                            float2* __restrict__ out) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] = make_float2(splatGrad[static_cast<size_t>(i) * 9], splatGrad[static_cast<size_t>(i) * 9 + 1]);
}

void Engine::enqueueBackward(FrameState& f, const SceneDev& s, const float* gradImage, float* gMeans,
                             float* gScales, float* gQuats, float* gOpacity, float* gSh) {
    if (f.n == 0) return;
    partial.reserve(std::max<size_t>(f.pairs, 1) * kPartialStride);
    splatGrad.reserve(f.n * 9);
    slotStart.reserve(f.n);
    const uint32_t n = static_cast<uint32_t>(f.n);
    const uint32_t grid = (n + 255) / 256;
    blockPairs.reserve(grid);
    // Pair slots in Gaussian index order.
    k_slotBlockSums<<<grid, 256, 0, stream>>>(n, f.tilesTouched.p, blockPairs.p, 0, 0);
    k_scanBlockSums<<<1, kScanBlock, 0, stream>>>(blockPairs.p, grid, 0);
    k_slotStart<<<grid, 256, 0, stream>>>(n, f.tilesTouched.p, blockPairs.p, slotStart.p, 0, 0);
    if (f.opts.tile_w == 8) launchBlendBackward<8, 8>(*this, f, gradImage);
    else if (f.opts.tile_w == 32) launchBlendBackward<32, 16>(*this, f, gradImage);
    else launchBlendBackward<16, 16>(*this, f, gradImage);
    k_gaussBackward<<<(n + kWave - 1) / kWave, kWave, 0, stream>>>(static_cast<int>(n), s.pos, s.scale, s.rot,
                                                                   s.opacity, s.sh, f.cp, f.tilesTouched.p,
                                                                   slotStart.p, partial.p, splatGrad.p, gMeans,
                                                                   gScales, gQuats, gOpacity, gSh, f.ctl.p, s.act);
    hipCheck(hipGetLastError(), "kernel launch");
}

}  // namespace ans

using namespace ans;


namespace {

template <class Fn>
int guardedDiff(Fn&& fn) {
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

}  // namespace

namespace ans {

SceneDev sceneOfParams(const ans_params* p) {
    if (!p) throw std::invalid_argument("params is null");
    if (p->n < 0 || p->n > (int64_t(1) << 30)) throw std::invalid_argument("Gaussian count must be in 0..2^30");
    if (p->sh_degree < 0 || p->sh_degree > 3) throw std::invalid_argument("sh_degree must be in 0..3");
    if (p->n > 0 && (!p->means || !p->log_scales || !p->quats || !p->opacity_logits || !p->sh))
        throw std::invalid_argument("parameter pointers must not be null");
    if (reinterpret_cast<uintptr_t>(p->quats) % 16 != 0 || reinterpret_cast<uintptr_t>(p->sh) % 16 != 0)
        throw std::invalid_argument("quats and sh must be 16-byte aligned");
    SceneDev s;
    s.n = p->n;
    s.shDegree = p->sh_degree;
    s.pos = p->means;
    s.scale = p->log_scales;
    s.rot = reinterpret_cast<const float4*>(p->quats);
    s.opacity = p->opacity_logits;
    s.sh = p->sh;
    s.shPlanar = false;
    s.raw = true;
    if (p->activated != 0 && p->activated != 1) throw std::invalid_argument("activated must be 0 or 1");
    s.act = p->activated != 0;
    return s;
}

}  // namespace ans

namespace {

SceneDev sceneOf(const ans_params* p) { return sceneOfParams(p); }

FrameState& frameOf(ans_diff* d, int64_t id) {
    if (!d) throw std::invalid_argument("handle is null");
    auto it = d->frames.find(id);
    if (it == d->frames.end()) throw std::invalid_argument("unknown or released frame");
    return *it->second;
}

}  // namespace

extern "C" {

int ans_diff_create(ans_diff** out) {
    return guardedDiff([&] {
        if (!out) throw std::invalid_argument("out is null");
        *out = nullptr;
        if (!ans_available()) throw std::runtime_error("no HIP device");
        auto d = std::make_unique<ans_diff>();
        d->engine = std::make_unique<Engine>();
        *out = d.release();
        return ANS_OK;
    });
}

void ans_diff_destroy(ans_diff* d) {
    if (!d) return;
    (void)hipStreamSynchronize(d->engine->stream);
    delete d;
}

int ans_diff_forward(ans_diff* d, const ans_params* params, const ans_camera* cam, const ans_options* options,
                     float* image, void* stream, int64_t* frame, ans_stats* stats) {
    return guardedDiff([&] {
        if (!d) throw std::invalid_argument("handle is null");
        checkCamera(cam);
        const SceneDev s = sceneOf(params);
        const ans_options o = resolveOptions(options);
        const CamParams cp = makeCamParams(*cam, o, s.shDegree);
        if (cp.tilesX > 65535 || cp.tilesY > 65535) throw std::invalid_argument("too many tiles");
        Engine& e = *d->engine;
        e.stream = static_cast<hipStream_t>(stream);  // NULL: the default stream, like any HIP call
        std::unique_ptr<FrameState> f;
        if (!d->pool.empty()) {
            f = std::move(d->pool.back());
            d->pool.pop_back();
        } else {
            f = std::make_unique<FrameState>();
        }
        f->savePixels = true;
        e.renderCounted(*f, s, cp, o, nullptr, image);
        e.fillStats(*f, stats);
        const int64_t id = d->next++;
        d->frames[id] = std::move(f);
        if (frame) *frame = id;
        return ANS_OK;
    });
}

int ans_diff_backward(ans_diff* d, int64_t frame, const ans_params* params, const float* grad_image,
                      const ans_grads* grads, void* stream) {
    return guardedDiff([&] {
        FrameState& f = frameOf(d, frame);
        const SceneDev s = sceneOf(params);
        if (static_cast<size_t>(s.n) != f.n) throw std::invalid_argument("parameter count differs from the forward");
        if (!grad_image || !grads || (s.n > 0 && (!grads->means || !grads->log_scales || !grads->quats ||
                                                  !grads->opacity_logits || !grads->sh)))
            throw std::invalid_argument("gradient pointers must not be null");
        if (reinterpret_cast<uintptr_t>(grads->quats) % 16 != 0)
            throw std::invalid_argument("grads quats must be 16-byte aligned");
        Engine& e = *d->engine;
        e.stream = static_cast<hipStream_t>(stream);  // NULL: the default stream, like any HIP call
        e.enqueueBackward(f, s, grad_image, grads->means, grads->log_scales, grads->quats, grads->opacity_logits,
                          grads->sh);
        return ANS_OK;
    });
}

int ans_diff_release(ans_diff* d, int64_t frame) {
    return guardedDiff([&] {
        if (!d) throw std::invalid_argument("handle is null");
        auto it = d->frames.find(frame);
        if (it != d->frames.end()) {
            d->pool.push_back(std::move(it->second));
            d->frames.erase(it);
            return ANS_OK;
        }
        auto vit = d->viewFrames.find(frame);
        if (vit == d->viewFrames.end()) throw std::invalid_argument("unknown or released frame");
        d->viewPool.push_back(std::move(vit->second));
        d->viewFrames.erase(vit);
        return ANS_OK;
    });
}

int ans_diff_splat_grads(ans_diff* d, int64_t frame, float* out) {
    return guardedDiff([&] {
        const FrameState& f = frameOf(d, frame);
        if (!out) throw std::invalid_argument("out is null");
        Engine& e = *d->engine;
        hipCheck(hipStreamSynchronize(e.stream), "hipStreamSynchronize");
        if (f.n && e.splatGrad.cap >= f.n * 9)
            hipCheck(hipMemcpy(out, e.splatGrad.p, f.n * 9 * sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy");
        return ANS_OK;
    });
}

int ans_diff_radii(ans_diff* d, int64_t frame, int32_t* radii, void* stream) {
    return guardedDiff([&] {
        const FrameState& f = frameOf(d, frame);
        if (!radii && f.n) throw std::invalid_argument("radii is null");
        const uint32_t n = static_cast<uint32_t>(f.n);
        if (n) {
            k_radii<<<(n + 255) / 256, 256, 0, static_cast<hipStream_t>(stream)>>>(n, f.rec.p, f.tilesTouched.p, radii);
            hipCheck(hipGetLastError(), "kernel launch");
        }
        return ANS_OK;
    });
}

int ans_diff_screen_grads(ans_diff* d, int64_t frame, float* grad_uv, void* stream) {
    return guardedDiff([&] {
        const FrameState& f = frameOf(d, frame);
        Engine& e = *d->engine;
        if (!grad_uv && f.n) throw std::invalid_argument("grad_uv is null");
        if (reinterpret_cast<uintptr_t>(grad_uv) % 8 != 0) throw std::invalid_argument("grad_uv must be 8-byte aligned");
        if (f.n && e.splatGrad.cap < f.n * 9) throw std::invalid_argument("no backward has run on this frame");
        const uint32_t n = static_cast<uint32_t>(f.n);
        if (n) {
            k_screenGrads<<<(n + 255) / 256, 256, 0, static_cast<hipStream_t>(stream)>>>(
                n, e.splatGrad.p, reinterpret_cast<float2*>(grad_uv));
            hipCheck(hipGetLastError(), "kernel launch");
        }
        return ANS_OK;
    });
}

int ans_diff_frame_counts(ans_diff* d, int64_t frame, int64_t* visible, int64_t* pairs, int64_t* tiles) {
    return guardedDiff([&] {
        const FrameState& f = frameOf(d, frame);
        if (visible) *visible = f.visible;
        if (pairs) *pairs = f.pairs;
        if (tiles) *tiles = static_cast<int64_t>(f.tiles);
        return ANS_OK;
    });
}

int ans_diff_frame_dump(ans_diff* d, int64_t frame, uint32_t* values, uint32_t* ranges) {
    return guardedDiff([&] {
        const FrameState& f = frameOf(d, frame);
        dumpFrame(d->engine->stream, f, nullptr, nullptr, nullptr, values, ranges);
        return ANS_OK;
    });
}

}  // extern "C"
