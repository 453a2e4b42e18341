// Kernels of the batched multi-view forward. Per-Gaussian arrays of view v sit at v * n, pair
// arrays at v * pairStride, tile ranges at v * maxTiles; a (view, tile) item is v * maxTiles + t.
// The per-view camera constants, output pointers and tile-grid sizes come from one device table
// uploaded per call, so a captured graph replays for any poses.
#pragma once

#include "backward.hpp"

namespace ans {


// Clears the first ctlCount control blocks (all but sticky) and rangeCount tile ranges.
[[maybe_unused]] static __global__ __launch_bounds__(256) void k_initViews(FrameCtl* __restrict__ ctl, uint32_t ctlCount,
                                                          uint2* __restrict__ ranges, uint32_t rangeCount) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < ctlCount) {
        ctl[i].epoch += 1;
        ctl[i].visible = 0;
        ctl[i].pairs = 0;
        ctl[i].overflow = 0;
    }
    if (i < rangeCount) ranges[i] = make_uint2(0, 0);
}

// Preprocess of every view, one lane per Gaussian: its parameters and SH (planar on the device,
// [coefficient * 3 + channel][n]) are loaded once into registers, its view-independent covariance
// is formed once per run of views with one splat scale (prepareGaussian), and it is projected
// into each view in turn (cameraTransform,
// finishProjection): the arithmetic of the single-view k_preprocess. Outputs for view v at
// v * n + i.
[[maybe_unused]] static __global__ __launch_bounds__(256) void k_preprocessViews(
    int n, int views, const float* __restrict__ pos, const float* __restrict__ scale, const float4* __restrict__ rot,
    const float* __restrict__ opacity, const float* __restrict__ sh, int shCoeffs, const CamParams* __restrict__ cams,
    SplatRec* __restrict__ rec, uint32_t* __restrict__ tilesTouched, uint32_t* __restrict__ depthKey,
    uint2* __restrict__ rect) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const F3 m = reinterpret_cast<const F3*>(pos)[i];
    const float op = opacity[i];
    const int K3 = shCoeffs * 3;
    alignas(16) float shr[48];
#pragma unroll
    for (int j = 0; j < 48; ++j) shr[j] = j < K3 ? sh[static_cast<size_t>(j) * n + i] : 0.0f;
    GaussProj g;
    float splatScale = cams[0].splatScale;
    prepareGaussian<false>(rot[i], reinterpret_cast<const F3*>(scale)[i], splatScale, g);
    for (int v = 0; v < views; ++v) {
        // A view with another splat scale (a reveal along a camera path) forms the covariance
        // again; the branch is uniform across the wavefront.
        if (cams[v].splatScale != splatScale) {
            splatScale = cams[v].splatScale;
            prepareGaussian<false>(rot[i], reinterpret_cast<const F3*>(scale)[i], splatScale, g);
        }
        const size_t o = static_cast<size_t>(v) * n + i;
        if (!cameraTransform(m, cams[v], g) || !finishProjection<false, false>(m, op, shr, 1, cams[v], g)) {
            storeCulled(o, rec, tilesTouched, depthKey, rect);
            continue;
        }
        storeSplat(g, o, rec, tilesTouched, depthKey, rect);
    }
}

// The differentiable version of k_preprocessViews: raw parameters ([n][coefficient][channel] SH,
// log-scales, logits, unnormalized quaternions; or activated scales and opacities with act), each
// view projected exactly as the single-view differentiable preprocess projects it
// (k_preprocess<false, true>), with the tile counts written for the adjoint's slot numbering.
[[maybe_unused]] static __global__ __launch_bounds__(256) void k_preprocessViewsRaw(
    int n, int views, const DiffParamsDev* __restrict__ prm, const CamParams* __restrict__ cams,
    SplatRec* __restrict__ rec, uint32_t* __restrict__ tilesTouched, uint32_t* __restrict__ depthKey,
    uint2* __restrict__ rect) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const DiffParamsDev P = *prm;
    const F3 m = reinterpret_cast<const F3*>(P.pos)[i];
    const float op = P.opacity[i];
    const int K3 = P.shCoeffs * 3;
    float shr[48];
    const float* src = P.sh + static_cast<size_t>(i) * K3;
#pragma unroll
    for (int j = 0; j < 48; ++j) shr[j] = j < K3 ? src[j] : 0.0f;
    GaussProj g;
    prepareGaussian<true>(P.rot[i], reinterpret_cast<const F3*>(P.scale)[i], cams[0].splatScale, g, P.act != 0);
    // The opacity is the same in every view: activated once here (as the single-view
    // finishProjection activates it), then passed on as an activated value.
    const float opAct = P.act ? op : 1.0f / (1.0f + expf(-op));
    for (int v = 0; v < views; ++v) {
        const size_t o = static_cast<size_t>(v) * n + i;
        if (!cameraTransform(m, cams[v], g) || !finishProjection<false, false>(m, opAct, shr, 1, cams[v], g)) {
            storeCulled(o, rec, tilesTouched, depthKey, rect);
            continue;
        }
        storeSplat(g, o, rec, tilesTouched, depthKey, rect);
    }
}

// Sort key of item v * maxTiles + t for longest-first order: 0x7fff - list length (lengths above
// 0x7fff count as 0x7fff); items past view v's tile count get 0xffff and come last.
[[maybe_unused]] static __global__ __launch_bounds__(256) void k_itemKeys(const uint2* __restrict__ ranges,
                                                         const int* __restrict__ tilesView, uint32_t maxTiles,
                                                         uint32_t items, uint32_t* __restrict__ keys) {
    const uint32_t it = blockIdx.x * 256 + threadIdx.x;
    if (it >= items) return;
    const uint32_t v = it / maxTiles, t = it - v * maxTiles;
    uint32_t key = 0xffffu;
    if (t < static_cast<uint32_t>(tilesView[v])) {
        const uint2 r = ranges[it];
        key = 0x7fffu - min(r.y - r.x, 0x7fffu);
    }
    keys[it] = key;
}

// Blend of every view's tiles in one launch, workgroups taken longest list first (itemOrder),
// each tile blended exactly as the single-view k_blend blends it. SAVE (the differentiable
// batch): float images outf[v] and per-pixel state at pixState + pixOff[v] instead of RGB8.
template <int TW, int TH, int BATCH, bool CULL, bool SAVE>
__global__ __launch_bounds__(TW* TH) void k_blendViews(const uint32_t* __restrict__ itemOrder, uint32_t maxTiles,
                                                       const uint2* __restrict__ ranges,
                                                       const uint32_t* __restrict__ sortedGauss, uint32_t pairStride,
                                                       const SplatRec* __restrict__ rec, uint32_t gaussStride,
                                                       const CamParams* __restrict__ cams, uint8_t* const* __restrict__ outs,
                                                       float* const* __restrict__ outf, uint2* __restrict__ pixState,
                                                       const uint64_t* __restrict__ pixOff,
                                                       const FrameCtl* __restrict__ ctl) {
    const uint32_t item = itemOrder[blockIdx.x];
    const uint32_t v = item / maxTiles, t = item - v * maxTiles;
    if (ctl[v].overflow) return;
    const int tilesX = cams[v].tilesX;
    const size_t go = static_cast<size_t>(v) * gaussStride;
    blendTile<TW, TH, BATCH, CULL, SAVE>(static_cast<int>(t) % tilesX, static_cast<int>(t) / tilesX, ranges[item],
                                         sortedGauss + static_cast<size_t>(v) * pairStride, rec + go, cams[v].width,
                                         cams[v].height, SAVE ? nullptr : outs[v], SAVE ? outf[v] : nullptr,
                                         SAVE ? pixState + pixOff[v] : nullptr);
}

// Adjoint of the differentiable batch's blend: every view's tiles in one launch, in the forward's
// longest-first order, each tile's adjoint exactly the single-view k_blendBackward's; view v's
// per-pair partial records start at pair pairBase[v] of the shared partial array.
template <int TW, int TH, int BATCH>
__global__ __launch_bounds__(TW* TH) void k_blendBackwardViews(
    const uint32_t* __restrict__ itemOrder, uint32_t maxTiles, const uint2* __restrict__ ranges,
    const uint32_t* __restrict__ sortedGauss, uint32_t pairStride, const SplatRec* __restrict__ rec,
    uint32_t gaussStride, const uint32_t* __restrict__ slotStart, const uint2* __restrict__ pixState,
    const uint64_t* __restrict__ pixOff, const CamParams* __restrict__ cams, const float* const* __restrict__ grads,
    const uint64_t* __restrict__ pairBase, float* __restrict__ partial) {
    const uint32_t item = itemOrder[blockIdx.x];
    const uint32_t v = item / maxTiles, t = item - v * maxTiles;
    const int tilesX = cams[v].tilesX;
    const size_t go = static_cast<size_t>(v) * gaussStride;
    blendBackwardTile<TW, TH, BATCH>(static_cast<int>(t) % tilesX, static_cast<int>(t) / tilesX, ranges[item],
                                     sortedGauss + static_cast<size_t>(v) * pairStride, rec + go, slotStart + go,
                                     pixState + pixOff[v], cams[v].width, cams[v].height, grads[v],
                                     partial + pairBase[v] * kPartialStride);
}

// Per-Gaussian adjoint of the differentiable batch: gaussBackwardWave over every view, the
// parameter and gradient arrays and the views' first pairs read from the call's device table.
[[maybe_unused]] static __global__ __launch_bounds__(kWave) void k_gaussBackwardViews(
    int n, int views, const DiffParamsDev* __restrict__ prm, const CamParams* __restrict__ cams,
    const uint32_t* __restrict__ tilesTouched, const uint32_t* __restrict__ slotStart,
    const uint64_t* __restrict__ pairBase, const float* __restrict__ partial) {
    gaussBackwardWave(n, views, *prm, cams, tilesTouched, slotStart, pairBase, partial);
}

}  // namespace ans
