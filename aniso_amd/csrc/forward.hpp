// Forward rasterizer kernels for gfx950: frame init, per-Gaussian preprocess, pair expansion,
// tile ranges and the tile blend. The library is compiled with -ffp-contract=off, so every
// a * b + c in the preprocess rounds twice, exactly like the CPU oracle; the blend spells out
// its fused multiply-adds explicitly.
#pragma once

#include "common.hpp"
#include "sort.hpp"

namespace ans {

// Clears the per-frame control block and the tile ranges, and bumps the epoch.
__global__ __launch_bounds__(256) void k_initFrame(FrameCtl* __restrict__ ctl, uint2* __restrict__ ranges,
                                                   uint32_t tiles) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i == 0) {
        ctl->epoch += 1;
        ctl->visible = 0;
        ctl->pairs = 0;
        ctl->overflow = 0;
    }
    constexpr uint32_t words = (sizeof(FrameCtl) - 32) / 4;
    if (i < words) reinterpret_cast<uint32_t*>(ctl)[8 + i] = 0;
    if (i < tiles) ranges[i] = make_uint2(0, 0);
}

struct F3 {
    float x, y, z;
};

// Real SH constants (src/sh.cpp:10-14; closed forms rounded to float).
__constant__ float kShC1 = 0.4886025119029199f;
__constant__ float kShC2[5] = {1.0925484305920792f, -1.0925484305920792f, 0.31539156525252005f, -1.0925484305920792f,
                               0.5462742152960396f};
__constant__ float kShC3[7] = {-0.5900435899266435f, 2.890611442640554f, -0.4570457994644658f, 0.3731763325901154f,
                               -0.4570457994644658f, 1.445305721320277f, -0.5900435899266435f};

// SH colour along a unit direction (src/sh.cpp:18-58, same basis order and operation order).
__device__ inline float3 shColor(const float* __restrict__ sh, int degree, float x, float y, float z) {
    float b[16];
    int count = 1;
    b[0] = 0.28209479177387814f;
    if (degree >= 1) {
        b[1] = -kShC1 * y;
        b[2] = kShC1 * z;
        b[3] = -kShC1 * x;
        count = 4;
    }
    if (degree >= 2) {
        const float xx = x * x, yy = y * y, zz = z * z;
        b[4] = kShC2[0] * x * y;
        b[5] = kShC2[1] * y * z;
        b[6] = kShC2[2] * (2 * zz - xx - yy);
        b[7] = kShC2[3] * x * z;
        b[8] = kShC2[4] * (xx - yy);
        count = 9;
        if (degree >= 3) {
            b[9] = kShC3[0] * y * (3 * xx - yy);
            b[10] = kShC3[1] * x * y * z;
            b[11] = kShC3[2] * y * (4 * zz - xx - yy);
            b[12] = kShC3[3] * z * (2 * zz - 3 * xx - 3 * yy);
            b[13] = kShC3[4] * x * (4 * zz - xx - yy);
            b[14] = kShC3[5] * z * (xx - yy);
            b[15] = kShC3[6] * x * (xx - 3 * yy);
            count = 16;
        }
    }
    float3 c = make_float3(0.5f, 0.5f, 0.5f);
    if (count == 16) {
        const float4* s4 = reinterpret_cast<const float4*>(sh);
        float v[48];
#pragma unroll
        for (int q = 0; q < 12; ++q) {
            const float4 t = s4[q];
            v[4 * q] = t.x; v[4 * q + 1] = t.y; v[4 * q + 2] = t.z; v[4 * q + 3] = t.w;
        }
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            c.x += b[k] * v[3 * k + 0];
            c.y += b[k] * v[3 * k + 1];
            c.z += b[k] * v[3 * k + 2];
        }
    } else {
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            if (k < count) {
                c.x += b[k] * sh[k * 3 + 0];
                c.y += b[k] * sh[k * 3 + 1];
                c.z += b[k] * sh[k * 3 + 2];
            }
        }
    }
    // std::max(0, v) semantics (NaN -> 0).
    c.x = (0.0f < c.x) ? c.x : 0.0f;
    c.y = (0.0f < c.y) ? c.y : 0.0f;
    c.z = (0.0f < c.z) ? c.z : 0.0f;
    return c;
}

__device__ inline float clampf(float v, float lo, float hi) { return (v < lo) ? lo : (hi < v) ? hi : v; }

// Per-Gaussian preprocess, one lane per Gaussian (the maths of src/render.cpp:39-104 with the
// FP32 camera transform of src/gpu_render.cu:94-96):
//   world -> camera; near cull; Sigma = M M^T, M = R(q) diag(s); EWA 2D covariance with the
//   clamped Jacobian and the 0.3 low-pass; conic; radius = ceil(3 sqrt(lambda_max)); centre in
//   pixels; screen cull; tile rect; SH colour.
// Writes the splat record {u, v, A, B | C, opacity, ext_x, ext_y | r, g, b, depth}, the packed
// tile rect, the tile count and the depth sort key (0xffffffff for culled Gaussians). ext_x/y are
// conservative half-extents of the region where alpha can reach 1/255, used by the blend to skip
// splats that cannot touch a wavefront's pixels (they do not affect any output).
__global__ __launch_bounds__(256) void k_preprocess(int n, const float* __restrict__ pos,
                                                    const float* __restrict__ scale, const float4* __restrict__ rot,
                                                    const float* __restrict__ opacity, const float* __restrict__ sh,
                                                    CamParams cam, float4* __restrict__ rec0,
                                                    float4* __restrict__ rec1, float4* __restrict__ rec2,
                                                    uint2* __restrict__ rect, uint32_t* __restrict__ tilesTouched,
                                                    uint32_t* __restrict__ depthKey, FrameCtl* __restrict__ ctl) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    bool vis = false;
    if (i < n) {
        uint32_t touched = 0;
        uint32_t key = 0xffffffffu;
        do {
            const F3 g = reinterpret_cast<const F3*>(pos)[i];
            const float px = cam.R[0] * g.x + cam.R[1] * g.y + cam.R[2] * g.z + cam.t[0];
            const float py = cam.R[3] * g.x + cam.R[4] * g.y + cam.R[5] * g.z + cam.t[1];
            const float pz = cam.R[6] * g.x + cam.R[7] * g.y + cam.R[8] * g.z + cam.t[2];
            if (!(pz >= 0.2f)) break;

            const float4 q = rot[i];  // w, x, y, z
            const float w = q.x, x = q.y, y = q.z, z = q.w;
            const float Rq[3][3] = {{1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)},
                                    {2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)},
                                    {2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)}};
            const F3 sc = reinterpret_cast<const F3*>(scale)[i];
            const float s[3] = {sc.x * cam.splatScale, sc.y * cam.splatScale, sc.z * cam.splatScale};
            float M[3][3], S[3][3];
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c) M[r][c] = Rq[r][c] * s[c];
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c) S[r][c] = M[r][0] * M[c][0] + M[r][1] * M[c][1] + M[r][2] * M[c][2];

            const float tx = clampf(px / pz, -cam.limX, cam.limX) * pz;
            const float ty = clampf(py / pz, -cam.limY, cam.limY) * pz;
            const float J[2][3] = {{cam.fx / pz, 0, -cam.fx * tx / (pz * pz)},
                                   {0, cam.fy / pz, -cam.fy * ty / (pz * pz)}};
            float T[2][3];
#pragma unroll
            for (int r = 0; r < 2; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c) T[r][c] = J[r][0] * cam.R[c] + J[r][1] * cam.R[3 + c] + J[r][2] * cam.R[6 + c];
            float TS[2][3];
#pragma unroll
            for (int r = 0; r < 2; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c) TS[r][c] = T[r][0] * S[0][c] + T[r][1] * S[1][c] + T[r][2] * S[2][c];
            const float a = TS[0][0] * T[0][0] + TS[0][1] * T[0][1] + TS[0][2] * T[0][2] + 0.3f;
            const float b = TS[0][0] * T[1][0] + TS[0][1] * T[1][1] + TS[0][2] * T[1][2];
            const float c = TS[1][0] * T[1][0] + TS[1][1] * T[1][1] + TS[1][2] * T[1][2] + 0.3f;
            const float det = a * c - b * b;
            if (!(det > 0.0f)) break;

            const float mid = 0.5f * (a + c);
            const float disc = mid * mid - det;
            const float lambda = mid + sqrtf((0.1f < disc) ? disc : 0.1f);
            const float rf = ceilf(3.0f * sqrtf(lambda));
            if (!(rf < 16777216.0f)) break;
            const int radius = static_cast<int>(rf);
            const float u = cam.fx * px / pz + cam.cx;
            const float v = cam.fy * py / pz + cam.cy;
            if (!isfinite(u) || !isfinite(v)) break;
            if (u + radius < 0 || v + radius < 0 || u - radius >= cam.width || v - radius >= cam.height) break;

            const int x0 = max(0, static_cast<int>(u - radius) / cam.tileW);
            const int y0 = max(0, static_cast<int>(v - radius) / cam.tileH);
            const int x1 = min(cam.tilesX - 1, static_cast<int>(u + radius) / cam.tileW);
            const int y1 = min(cam.tilesY - 1, static_cast<int>(v + radius) / cam.tileH);

            const float d0 = g.x - cam.eye[0], d1 = g.y - cam.eye[1], d2 = g.z - cam.eye[2];
            const float len = sqrtf(d0 * d0 + d1 * d1 + d2 * d2);
            const float3 rgb = shColor(sh + static_cast<size_t>(i) * cam.shCoeffs * 3, cam.shDegree, d0 / len,
                                       d1 / len, d2 / len);
            const float op = opacity[i];

            // Region where op * exp(-q/2) >= 1/255: q <= 2 ln(255 op); half-extents
            // sqrt(k * a), sqrt(k * c) of that ellipse, widened by a 5% + 0.1 margin that covers
            // the FP32 rounding of the blend's exponent by orders of magnitude.
            float ex = -1e30f, ey = -1e30f;
            if (op >= 1.0f / 255.0f) {
                const float k = fmaxf(2.0f * logf(255.0f * op), 0.0f) * 1.05f + 0.1f;
                ex = sqrtf(k * a) + 0.01f;
                ey = sqrtf(k * c) + 0.01f;
            }
            rec0[i] = make_float4(u, v, c / det, -b / det);
            rec1[i] = make_float4(a / det, op, ex, ey);
            rec2[i] = make_float4(rgb.x, rgb.y, rgb.z, pz);
            rect[i] = make_uint2(static_cast<uint32_t>(x0) | (static_cast<uint32_t>(y0) << 16),
                                 static_cast<uint32_t>(x1) | (static_cast<uint32_t>(y1) << 16));
            touched = static_cast<uint32_t>((x1 - x0 + 1) * (y1 - y0 + 1));
            key = __float_as_uint(pz);
            vis = true;
        } while (false);
        tilesTouched[i] = touched;
        depthKey[i] = key;
    }
    const uint64_t m = __ballot(vis);
    if (__lane_id() == 0 && m) atomicAdd(&ctl->visible, static_cast<uint32_t>(__popcll(m)));
}

constexpr int kDupBlock = 256;
constexpr int kDupItems = 4;
constexpr int kDupTile = kDupBlock * kDupItems;

// Pair expansion in depth order. Ranks [part * 1024, +1024) of the depth-sorted Gaussians get
// their pair offsets by a block scan plus decoupled look-back (the prefix sum of
// gpu_render.cu:324-327 without the host read-back of :329); the block then writes its pairs
// cooperatively, one pair per lane per step, each Gaussian's tiles in row-major order
// (gpu_render.cu:155-156). Pair = (tile id, Gaussian index). pairStart[g] records each
// Gaussian's first pair slot. Digit histograms of the tile ids feed the pair sort.
// A pair count above `cap` sets ctl->overflow and writes nothing past the capacity.
__global__ __launch_bounds__(kDupBlock) void k_expandPairs(
    const uint32_t* __restrict__ sortedIdx, const uint32_t* __restrict__ tilesTouched, const uint2* __restrict__ rect,
    uint32_t n, uint32_t cap, int tilesX, int digitBits, int passes, uint32_t* __restrict__ pairStart,
    uint32_t* __restrict__ tileOut, uint32_t* __restrict__ gaussOut, uint32_t* __restrict__ partCounter,
    uint64_t* __restrict__ status, FrameCtl* __restrict__ ctl, uint32_t pass) {
    __shared__ uint32_t sIncl[kDupTile];
    __shared__ uint32_t sGauss[kDupTile];
    __shared__ uint2 sRect[kDupTile];
    __shared__ uint32_t sHist[3 * kRadix];
    __shared__ uint32_t sScan[kDupBlock / kWave];
    __shared__ uint32_t sPart, sBase;

    const int tid = threadIdx.x;
    if (tid == 0) sPart = atomicAdd(partCounter, 1u);
    for (int i = tid; i < 3 * kRadix; i += kDupBlock) sHist[i] = 0;
    __syncthreads();
    const uint32_t part = sPart;
    const uint32_t base = part * kDupTile;

    uint32_t cnt[kDupItems];
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < kDupItems; ++k) {
        const uint32_t r = base + tid * kDupItems + k;
        uint32_t c = 0, g = 0;
        if (r < n) {
            g = sortedIdx[r];
            c = tilesTouched[g];
        }
        sGauss[tid * kDupItems + k] = g;
        cnt[k] = c;
        sum += c;
    }
    uint32_t total;
    const uint32_t excl = blockExclusiveScan<kDupBlock>(sum, sScan, &total);
    if (tid == 0) {
        const uint32_t tag = ctl->epoch * 16u + pass;
        uint64_t* st = status + part;
        uint32_t earlier = 0;
        if (part == 0) {
            storeStatus(st, statusWord(tag, kFlagPrefix, total));
        } else {
            storeStatus(st, statusWord(tag, kFlagAggregate, total));
            earlier = lookBack(status, part, 1, 0, tag);
            storeStatus(st, statusWord(tag, kFlagPrefix, earlier + total));
        }
        sBase = earlier;
        if (part == gridDim.x - 1) {
            const uint64_t all = static_cast<uint64_t>(earlier) + total;
            ctl->pairs = static_cast<uint32_t>(all > 0xffffffffull ? 0xffffffffull : all);
            if (all > cap) {
                ctl->overflow = 1;
                ctl->sticky = 1;
            }
        }
    }
    uint32_t run = excl;
#pragma unroll
    for (int k = 0; k < kDupItems; ++k) {
        run += cnt[k];
        sIncl[tid * kDupItems + k] = run;
    }
    __syncthreads();
    const uint32_t gbase = sBase;
#pragma unroll
    for (int k = 0; k < kDupItems; ++k) {
        const uint32_t r = base + tid * kDupItems + k;
        const uint32_t j = tid * kDupItems + k;
        if (r < n && cnt[k] > 0) {
            const uint32_t g = sGauss[j];
            pairStart[g] = gbase + sIncl[j] - cnt[k];
            sRect[j] = rect[g];
        }
    }
    __syncthreads();
    if (static_cast<uint64_t>(gbase) + total > cap) return;

    const uint32_t dmask = (1u << digitBits) - 1u;
    for (uint32_t s = tid; s < total; s += kDupBlock) {
        // First local rank whose inclusive end exceeds s.
        uint32_t lo = 0, hi = kDupTile - 1;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (sIncl[mid] > s) hi = mid; else lo = mid + 1;
        }
        const uint32_t start = lo ? sIncl[lo - 1] : 0u;
        const uint32_t k = s - start;
        const uint2 rp = sRect[lo];
        const uint32_t x0 = rp.x & 0xffffu, y0 = rp.x >> 16, x1 = rp.y & 0xffffu;
        const uint32_t w = x1 - x0 + 1;
        const uint32_t ty = y0 + k / w, tx = x0 + k % w;
        const uint32_t tile = ty * static_cast<uint32_t>(tilesX) + tx;
        tileOut[gbase + s] = tile;
        gaussOut[gbase + s] = sGauss[lo];
        for (int p = 0; p < passes; ++p) atomicAdd(&sHist[p * kRadix + ((tile >> (p * digitBits)) & dmask)], 1u);
    }
    __syncthreads();
    for (int i = tid; i < passes * kRadix; i += kDupBlock)
        if (sHist[i]) atomicAdd(&ctl->pairHist[0][0] + i, sHist[i]);
}

// Start and end of each tile's run in the sorted pairs (gpu_render.cu:164-177); ranges were
// cleared to [0, 0) by the frame init.
__global__ __launch_bounds__(256) void k_tileRanges(const uint32_t* __restrict__ sortedTile,
                                                    const FrameCtl* __restrict__ ctl, uint2* __restrict__ ranges) {
    if (ctl->overflow) return;
    const uint32_t total = ctl->pairs;
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    if (p >= total) return;
    const uint32_t t = sortedTile[p];
    if (p == 0 || sortedTile[p - 1] != t) ranges[t].x = p;
    if (p == total - 1 || sortedTile[p + 1] != t) ranges[t].y = p + 1;
}

// Front-to-back blend, one workgroup per TW x TH tile, one lane per pixel; each wavefront owns
// an 8 x 8 block of the tile (src/gpu_render.cu:180-233 and src/render.cpp:193-221 define the
// per-pixel rule: skip power > 0, alpha = min(0.99, o exp(power)), skip alpha < 1/255, stop once
// T (1 - alpha) < 1e-4, C += rgb alpha T). BATCH splat records are staged in LDS per round;
// a wavefront skips records whose 1/255 extent misses its 8 x 8 block, stops when all of its
// pixels are done, and the workgroup stops when all of its wavefronts are.
template <int TW, int TH, int BATCH>
__global__ __launch_bounds__(TW* TH) void k_blend(const uint2* __restrict__ ranges,
                                                  const uint32_t* __restrict__ sortedGauss,
                                                  const float4* __restrict__ rec0, const float4* __restrict__ rec1,
                                                  const float4* __restrict__ rec2, int width, int height, int tilesX,
                                                  uint8_t* __restrict__ out8, float* __restrict__ outf,
                                                  const FrameCtl* __restrict__ ctl) {
    constexpr int kThreads = TW * TH;
    constexpr int kWavesX = TW / 8;
    __shared__ float4 s0[BATCH], s1[BATCH], s2[BATCH];

    if (ctl->overflow) return;
    const int tile = blockIdx.x;
    const int tileX = tile % tilesX, tileY = tile / tilesX;
    const int tid = threadIdx.x;
    const int wave = tid / kWave, lane = __lane_id();
    const int bx = tileX * TW + (wave % kWavesX) * 8, by = tileY * TH + (wave / kWavesX) * 8;
    const int x = bx + (lane & 7), y = by + (lane >> 3);
    const bool inside = x < width && y < height;
    const float px = x + 0.5f, py = y + 0.5f;
    // Pixel-centre bounds of this wavefront's 8 x 8 block.
    const float bx0 = bx + 0.5f, bx1 = bx + 7.5f, by0 = by + 0.5f, by1 = by + 7.5f;

    const uint2 range = ranges[tile];
    float T = 1.0f;
    float cr = 0.0f, cg = 0.0f, cb = 0.0f;
    bool done = !inside;
    for (uint32_t start = range.x; start < range.y; start += BATCH) {
        if (__syncthreads_count(done) == kThreads) break;
        const int cnt = min(static_cast<uint32_t>(BATCH), range.y - start);
        for (int k = tid; k < cnt; k += kThreads) {
            const uint32_t g = sortedGauss[start + k];
            s0[k] = rec0[g];
            s1[k] = rec1[g];
            s2[k] = rec2[g];
        }
        __syncthreads();
        if (__ballot(!done) != 0) {
            for (int j = 0; j < cnt; ++j) {
                const float4 a = s0[j];
                const float4 b = s1[j];
                if (a.x + b.z < bx0 || a.x - b.z > bx1 || a.y + b.w < by0 || a.y - b.w > by1) continue;
                if (!done) {
                    const float dx = px - a.x, dy = py - a.y;
                    const float power = -0.5f * (a.z * dx * dx + b.x * dy * dy) - a.w * dx * dy;
                    if (!(power > 0.0f)) {
                        const float alpha = fminf(0.99f, b.y * __expf(power));
                        if (alpha >= 1.0f / 255.0f) {
                            const float nextT = T * (1.0f - alpha);
                            if (nextT < 1e-4f) {
                                done = true;
                            } else {
                                const float4 c = s2[j];
                                const float w = alpha * T;
                                cr = __builtin_fmaf(c.x, w, cr);
                                cg = __builtin_fmaf(c.y, w, cg);
                                cb = __builtin_fmaf(c.z, w, cb);
                                T = nextT;
                            }
                        }
                    }
                }
                if ((j & 7) == 7 && __ballot(!done) == 0) break;
            }
        }
    }
    if (inside) {
        const size_t p = static_cast<size_t>(y) * width + x;
        if (out8) {
            out8[3 * p + 0] = static_cast<uint8_t>(fminf(fmaxf(cr, 0.f), 1.f) * 255.f + 0.5f);
            out8[3 * p + 1] = static_cast<uint8_t>(fminf(fmaxf(cg, 0.f), 1.f) * 255.f + 0.5f);
            out8[3 * p + 2] = static_cast<uint8_t>(fminf(fmaxf(cb, 0.f), 1.f) * 255.f + 0.5f);
        }
        if (outf) {
            outf[3 * p + 0] = cr;
            outf[3 * p + 1] = cg;
            outf[3 * p + 2] = cb;
        }
    }
}

}  // namespace ans
