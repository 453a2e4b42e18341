// Shared device-side definitions for the gfx950 rasterizer kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ans {

constexpr int kWave = 64;  // CDNA wavefront width

// Per-frame camera constants in FP32, derived once on the host from the double camera exactly as
// the CPU renderer derives its own (src/render.cpp:58-59, src/camera.cpp:19-23,
// src/gpu_render.cu:301-316). Passed by value as a kernel argument.
struct CamParams {
    float R[9];  // world to camera, row major
    float t[3];
    float eye[3];  // camera centre in world space
    float fx, fy, cx, cy;
    float limX, limY;  // 1.3 * tan(half field of view), the Jacobian clamp
    int width, height;
    int tileW, tileH, tilesX, tilesY;
    int tileShiftX, tileShiftY;  // log2 of tileW, tileH (powers of two)
    int shDegree;  // degree evaluated (option clamped to the scene)
    int shCoeffs;  // coefficients stored per channel in the scene
    float splatScale;
};

// One Gaussian's splat record: 64 bytes, one aligned cache line, so that the per-pair gathers of
// the blend, its adjoint and the pair expansion each touch a single line.
struct alignas(64) SplatRec {
    float4 geo;   // u, v (pixels), conic A, B
    float4 geo2;  // conic C, opacity, 0, 0 (the blend culls from the conic and opacity)
    float4 col;   // r, g, b, depth
    uint4 meta;   // tile rect (x0 | y0 << 16, x1 | y1 << 16), tiles touched (0: culled), radius (px)
};

// Small per-frame control block in device memory. The init kernel clears it; kernels publish
// counts into it; the host reads it back with the image.
struct FrameCtl {
    uint32_t epoch;       // bumped once per frame
    uint32_t visible;     // Gaussians that survived culling
    uint32_t pairs;       // Gaussian-tile pairs (sum of tiles touched)
    uint32_t overflow;    // 1 if pairs exceeded the pair buffers' capacity
    uint32_t sticky;      // overflow of any frame since the host last cleared it (not cleared by init)
    uint32_t finished;    // blend workgroups that have ended (stage timing)
    uint32_t sortCount;   // keys kept by a depth-sort pass that dropped culled Gaussians
    uint32_t pad;
};

// A frame's pair count as the pair-count scan publishes it to pinned host memory: the fields,
// then seq (the host's number for the frame) with a system-scope release, so a host that sees
// seq sees the fields. The host waits on seq instead of synchronizing the stream, and the kernels
// it enqueued after the count keep the GPU busy meanwhile.
struct CountMirror {
    uint32_t seq;
    uint32_t pairs;
    uint32_t visible;
    uint32_t overflow;
    // Frames on this stream whose pairs exceeded the buffers and that the host has not handled
    // yet, and the largest pair count among them (the buffers must grow to that, not to the
    // count of whichever frame ran last). The scan increments them; the host clears them.
    uint32_t sticky;
    uint32_t stickyPairs;
};

// Stage timestamps of one frame, read from the GPU's constant-rate wall clock (s_memrealtime) by
// the kernels themselves and stored straight to pinned host memory: [0] preprocess start, [1]
// pair counting start (end of the depth sort), [2] blend start, [3] blend end. Stamps taken in
// kernels cost nothing between launches; events recorded between two kernels separate them by
// several microseconds.
struct FrameStamps {
    uint64_t t[4];
};

// Orders LDS accesses of different lanes of one wavefront: writes before the call are visible
// to reads after it, and the compiler may not move or forward memory operations across it.
__device__ inline void waveSync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline uint64_t lanesBelow() {
    const uint32_t lane = __lane_id();
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// DPP lane moves within a 64-lane wavefront (GFX9 controls): row_shr:n gives lane l of each
// 16-lane row lane l - n of the same row; row_bcast:15 gives every lane of rows 1 and 3 (row mask
// 0xa) lane 15 of the row below; row_bcast:31 gives rows 2 and 3 (row mask 0xc) lane 31;
// wave_shl:1 gives lane l lane l + 1, wave_shr:1 lane l - 1. Lanes without a source, or in masked
// rows, get `old`.
// These run on the vector ALU; lane shuffles through the LDS crossbar (ds_bpermute) cost a
// round trip each.
enum : int {
    kDppRowShr1 = 0x111,
    kDppRowShr2 = 0x112,
    kDppRowShr4 = 0x114,
    kDppRowShr8 = 0x118,
    kDppWaveShl1 = 0x130,
    kDppWaveShr1 = 0x138,
    kDppRowBcast15 = 0x142,
    kDppRowBcast31 = 0x143,
};

template <int CTRL, int ROWS = 0xf>
__device__ inline int dppMove(int old, int v) {
    return __builtin_amdgcn_update_dpp(old, v, CTRL, ROWS, 0xf, false);
}

// Inclusive scan over the 64 lanes with an associative integer operation whose identity is
// ident: steps of 1, 2, 4 and 8 lanes within each row, then row 0's total into row 1 and row 2's
// into row 3, then rows 0-1's total into rows 2-3.
template <typename Op>
__device__ inline int waveInclusive(int v, int ident, Op op) {
    v = op(v, dppMove<kDppRowShr1>(ident, v));
    v = op(v, dppMove<kDppRowShr2>(ident, v));
    v = op(v, dppMove<kDppRowShr4>(ident, v));
    v = op(v, dppMove<kDppRowShr8>(ident, v));
    v = op(v, dppMove<kDppRowBcast15, 0xa>(ident, v));
    v = op(v, dppMove<kDppRowBcast31, 0xc>(ident, v));
    return v;
}

}  // namespace ans
