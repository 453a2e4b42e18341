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
    int shDegree;  // degree evaluated (option clamped to the scene)
    int shCoeffs;  // coefficients stored per channel in the scene
    float splatScale;
};

// Small per-frame control block in device memory. The init kernel clears it; kernels publish
// counts into it; the host reads it back with the image.
struct FrameCtl {
    uint32_t epoch;       // bumped once per frame; tags look-back status words
    uint32_t visible;     // Gaussians that survived culling
    uint32_t pairs;       // Gaussian-tile pairs (sum of tiles touched)
    uint32_t overflow;    // 1 if pairs exceeded the pair buffers' capacity
    uint32_t sticky;      // overflow of any frame since the host last cleared it (not cleared by init)
    uint32_t pad[3];
    uint32_t partCounter[8];  // virtual block ids for the look-back kernels, one per pass
    uint32_t depthHist[4][256];  // digit histograms of the depth keys
    uint32_t pairHist[3][256];   // digit histograms of the pair tile ids
};

// Look-back status word: [63:32] tag (epoch * 16 + pass), [31:30] flag, [29:0] count.
constexpr uint64_t kFlagAggregate = 1ull << 30;
constexpr uint64_t kFlagPrefix = 2ull << 30;
constexpr uint32_t kCountMask = (1u << 30) - 1;

__device__ inline uint64_t statusWord(uint32_t tag, uint64_t flag, uint32_t count) {
    return (static_cast<uint64_t>(tag) << 32) | flag | count;
}

__device__ inline void storeStatus(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline uint64_t loadStatus(const uint64_t* p) {
    return __hip_atomic_load(const_cast<uint64_t*>(p), __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}

// Sum of the counts of all earlier partitions for one bin: walks back from `part - 1`, adding
// aggregates until it meets an inclusive prefix (decoupled look-back). Earlier partitions were
// claimed by blocks that are already running, so every wait ends.
__device__ inline uint32_t lookBack(const uint64_t* status, uint32_t part, uint32_t stride, uint32_t bin,
                                    uint32_t tag) {
    uint32_t sum = 0;
    for (int64_t p = static_cast<int64_t>(part) - 1; p >= 0; --p) {
        const uint64_t* w = status + static_cast<uint64_t>(p) * stride + bin;
        uint64_t v;
        while (true) {
            v = loadStatus(w);
            if (static_cast<uint32_t>(v >> 32) == tag && (v & (3ull << 30)) != 0) break;
            __builtin_amdgcn_s_sleep(1);
        }
        sum += static_cast<uint32_t>(v) & kCountMask;
        if (v & kFlagPrefix) break;
    }
    return sum;
}

__device__ inline uint64_t lanesBelow() {
    const uint32_t lane = __lane_id();
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

}  // namespace ans
