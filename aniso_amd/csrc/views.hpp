// Kernels of the batched multi-view forward. Per-Gaussian arrays of view v sit at v * n, pair
// arrays at v * pairStride, tile ranges at v * maxTiles; a (view, tile) item is v * maxTiles + t.
// The per-view camera constants, output pointers and tile-grid sizes come from one device table
// uploaded per call, so a captured graph replays for any poses.
#pragma once

#include "backward.hpp"

namespace ans {

// Parameter and gradient arrays of a batched differentiable call, read by the kernels from the
// per-call device table (so that a captured graph replays for any tensors): the layout of
// ans_params / ans_grads (means [n][3], log-scales [n][3], quaternions [n][4], opacity logits [n],
// SH [n][coefficient][channel]).
struct DiffParamsDev {
    const float* pos;
    const float* scale;
    const float4* rot;
    const float* opacity;
    const float* sh;
    float* gMeans;
    float* gScales;
    float* gQuats;
    float* gOpacity;
    float* gSh;
    float2* screenGrads;  // [view][n] dL/d(u, v), or null
    int shCoeffs;
    int act;  // ans_params.activated
};

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
    SplatRec* __restrict__ rec, uint32_t* __restrict__ tilesTouched, uint32_t* __restrict__ depthKey) {
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
            storeCulled(o, rec, tilesTouched, depthKey);
            continue;
        }
        storeSplat(g, o, rec, tilesTouched, depthKey);
    }
}

// The differentiable version of k_preprocessViews: raw parameters ([n][coefficient][channel] SH,
// log-scales, logits, unnormalized quaternions; or activated scales and opacities with act), each
// view projected exactly as the single-view differentiable preprocess projects it
// (k_preprocess<false, true>), with the tile counts written for the adjoint's slot numbering.
[[maybe_unused]] static __global__ __launch_bounds__(256) void k_preprocessViewsRaw(
    int n, int views, const DiffParamsDev* __restrict__ prm, const CamParams* __restrict__ cams,
    SplatRec* __restrict__ rec, uint32_t* __restrict__ tilesTouched, uint32_t* __restrict__ depthKey) {
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
    for (int v = 0; v < views; ++v) {
        const size_t o = static_cast<size_t>(v) * n + i;
        if (!cameraTransform(m, cams[v], g) || !finishProjection<false, true>(m, op, shr, 1, cams[v], g)) {
            storeCulled(o, rec, tilesTouched, depthKey);
            continue;
        }
        storeSplat(g, o, rec, tilesTouched, depthKey);
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

// Per-Gaussian adjoint of the differentiable batch, one wavefront per 64 consecutive Gaussians:
// for each view in turn (v = 0, 1, ...) the wavefront streams its pair records of that view
// (one contiguous run, as in k_gaussBackward), each lane sums its Gaussian's records in slot
// order, projects the Gaussian into the view again and applies the view-dependent part of the
// adjoint (gaussAdjointView); the mean and SH gradients, dL/dSigma and dL/dopacity are added in
// view order into running sums (SH sums in LDS), and the view-independent part
// (gaussAdjointShape, linear in dL/dSigma and dL/dopacity) runs once on the sums. Each
// gradient is the sum over views of the single-view gradients (up to the rounding of that
// linear map), without atomics. Culled (view, Gaussian) pairs contribute nothing. With
// P.screenGrads, dL/d(u, v) per view.
[[maybe_unused]] static __global__ __launch_bounds__(kWave) void k_gaussBackwardViews(
    int n, int views, const DiffParamsDev* __restrict__ prm, const CamParams* __restrict__ cams,
    const uint32_t* __restrict__ tilesTouched, const uint32_t* __restrict__ slotStart,
    const uint64_t* __restrict__ pairBase, const float* __restrict__ partial) {
    // Only the SH gradient sums live in LDS (12.5 KB per wavefront, so that LDS allows ten
    // wavefronts per CU); each lane reads its Gaussian's SH from global memory, where its
    // 48-float row stays in cache across the views.
    __shared__ float4 sRec[kWave][3];
    __shared__ float sGsh[kWave * kShRow];
    const DiffParamsDev P = *prm;
    const int lane = threadIdx.x;
    const int g0 = blockIdx.x * kWave;
    const int rows = min(kWave, n - g0);
    const int i = g0 + lane;
    const bool valid = lane < rows;
    const int K3 = P.shCoeffs * 3;  // at most 48
    const float* shRow = P.sh + static_cast<size_t>(valid ? i : g0) * K3;
    float* gshRow = sGsh + lane * kShRow;
    for (int j = 0; j < K3; ++j) gshRow[j] = 0.0f;
    F3 m{0.0f, 0.0f, 0.0f};
    float logit = 0.0f;
    GaussProj g;
    if (valid) {
        m = reinterpret_cast<const F3*>(P.pos)[i];
        logit = P.opacity[i];
        prepareGaussian<true>(P.rot[i], reinterpret_cast<const F3*>(P.scale)[i], cams[0].splatScale, g, P.act != 0);
    }
    float gm[3] = {0, 0, 0}, gSig[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, go = 0.0f;
    const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (int v = 0; v < views; ++v) {
        const size_t vo = static_cast<size_t>(v) * n;
        const uint32_t tt = valid ? tilesTouched[vo + i] : 0u;
        const uint32_t s0 = valid ? slotStart[vo + i] : 0u;
        const uint32_t r0 = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(s0)));
        const uint32_t r1 = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(s0 + tt), rows - 1));
        const float4* rec = reinterpret_cast<const float4*>(partial + pairBase[v] * kPartialStride);
        float4 a0 = zero, a1 = zero, a2 = zero;
        if (r0 + lane < r1) {
            const size_t r = static_cast<size_t>(r0 + lane) * 3;
            a0 = rec[r];
            a1 = rec[r + 1];
            a2 = rec[r + 2];
        }
        float gs[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (uint32_t c0 = r0; c0 < r1; c0 += kWave) {
            waveSync();
            sRec[lane][0] = a0;
            sRec[lane][1] = a1;
            sRec[lane][2] = a2;
            const uint32_t nx = c0 + kWave + lane;
            if (nx < r1) {
                const size_t r = static_cast<size_t>(nx) * 3;
                a0 = rec[r];
                a1 = rec[r + 1];
                a2 = rec[r + 2];
            }
            waveSync();
            const uint32_t a = max(s0, c0), b = min(s0 + tt, c0 + kWave);
            for (uint32_t k = a; k < b; ++k) {
                const float4 x = sRec[k - c0][0], y = sRec[k - c0][1], z = sRec[k - c0][2];
                gs[0] += x.x; gs[1] += x.y; gs[2] += x.z; gs[3] += x.w;
                gs[4] += y.x; gs[5] += y.y; gs[6] += y.z; gs[7] += y.w;
                gs[8] += z.x;
            }
        }
        if (valid && P.screenGrads) P.screenGrads[vo + i] = make_float2(gs[0], gs[1]);
        if (tt > 0 && cameraTransform(m, cams[v], g) && finishProjection<false, true>(m, logit, shRow, 1, cams[v], g)) {
            float vm[3], vSig[3][3];
            gaussAdjointView<true>(g, cams[v], gs, shRow, vm, vSig, gshRow);
            for (int j = 0; j < 3; ++j) gm[j] += vm[j];
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) gSig[r][c] += vSig[r][c];
            go += gs[5];
        }
    }
    if (valid) {
        // The view-independent part once, on the sums; the rotation and scale matrices are formed
        // again here so that they need not stay in registers through the view loop.
        GaussProj h;
        prepareGaussian<true>(P.rot[i], reinterpret_cast<const F3*>(P.scale)[i], cams[0].splatScale, h, P.act != 0);
        h.opacity = h.act ? logit : 1.0f / (1.0f + expf(-logit));  // as finishProjection sets it
        float gls[3], gq[4], gl;
        gaussAdjointShape(h, cams[0].splatScale, gSig, go, gls, gq, &gl);
        for (int j = 0; j < 3; ++j) P.gMeans[3 * i + j] = gm[j], P.gScales[3 * i + j] = gls[j];
        reinterpret_cast<float4*>(P.gQuats)[i] = make_float4(gq[0], gq[1], gq[2], gq[3]);
        P.gOpacity[i] = gl;
    }
    waveSync();
    float* dst = P.gSh + static_cast<size_t>(g0) * K3;
    const int total = rows * K3;
    const float rcpK3 = 1.0f / static_cast<float>(K3);
#pragma unroll
    for (int j = 0; j < 48; ++j) {
        const int e = lane + j * kWave;
        if (e < total) {
            const int r = rowOf(e, rcpK3);
            dst[e] = sGsh[r * kShRow + (e - r * K3)];
        }
    }
}

}  // namespace ans
