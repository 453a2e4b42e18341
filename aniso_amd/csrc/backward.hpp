// Backward kernels for gfx950: the adjoint of the tile blend and of the per-Gaussian projection.
// There is no backward in the reference (its training uses a third-party rasterizer); the
// forward these differentiate is the reference's (src/render.cpp:39-104, :193-221).
//
// Gradients are reduced without atomics: every (tile, Gaussian) pair is owned by one workgroup,
// which sums its pixels' contributions in a fixed order and writes one 9-float partial per pair
// into the pair's slot in the unsorted depth-rank order; each Gaussian's slots are contiguous
// ([pairStart, pairStart + tiles touched)), so one lane per Gaussian sums them in order.
#pragma once

#include "forward.hpp"

namespace ans {

template <int CTRL, int ROW_MASK, int BANK_MASK>
__device__ inline float dppAdd(float v) {
    const int moved = __builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, BANK_MASK, false);
    return v + __int_as_float(moved);
}

// Sum over the 64 lanes, valid in lane 63: inclusive scan within rows of 16 (row_shr 1, 2, 4, 8),
// then row_bcast:15 and row_bcast:31 carry the row totals up.
__device__ inline float waveSumLast(float v) {
    v = dppAdd<0x111, 0xf, 0xf>(v);
    v = dppAdd<0x112, 0xf, 0xf>(v);
    v = dppAdd<0x114, 0xf, 0xe>(v);
    v = dppAdd<0x118, 0xf, 0xc>(v);
    v = dppAdd<0x142, 0xa, 0xf>(v);
    v = dppAdd<0x143, 0xc, 0xf>(v);
    return v;
}

constexpr int kMaxCheckpoints = 16;

// Adjoint of k_blend for one tile. Pass 1 re-runs the front-to-back walk to find each pixel's last
// contributing splat and its final transmittance, recording the transmittance at the start of
// every ckEvery-th batch in LDS. Pass 2 walks the batches back to front: within a batch,
// T_i = T_{i+1} / (1 - alpha_i) (the same alpha the forward computed, so 1 - alpha is the same
// float), and at each checkpointed batch start T is reset to the recorded forward value, so
// rounding cannot accumulate across batches. Per contributing splat i of a pixel with image
// gradient g = dL/dC:
//   dL/drgb_i = g alpha_i T_i,  dL/dalpha_i = T_i g . (rgb_i - accum),  accum <- alpha_i rgb_i +
//   (1 - alpha_i) accum  (accum = colour behind i per unit of the transmittance after i);
//   alpha = o G (no gradient where the 0.99 clamp is active), G = exp(power),
//   power = -(A dx^2 + C dy^2) / 2 - B dx dy, dx = px - u, dy = py - v.
// partial[slot * 9 + {u, v, A, B, C, opacity, r, g, b}].
template <int TW, int TH, int BATCH>
__global__ __launch_bounds__(TW* TH) void k_blendBackward(
    const uint2* __restrict__ ranges, const uint32_t* __restrict__ sortedGauss, const float4* __restrict__ rec0,
    const float4* __restrict__ rec1, const float4* __restrict__ rec2, const uint2* __restrict__ rect,
    const uint32_t* __restrict__ pairStart, int width, int height, int tilesX, const float* __restrict__ gradImage,
    float* __restrict__ partial, const FrameCtl* __restrict__ ctl) {
    constexpr int kThreads = TW * TH;
    constexpr int kWaves = kThreads / kWave;
    constexpr int kWavesX = TW / 8;
    __shared__ float4 s0[BATCH], s1[BATCH], s2[BATCH];
    __shared__ uint32_t sGauss[BATCH];
    __shared__ uint8_t sMask[BATCH];
    __shared__ uint8_t sList[kWaves][BATCH];
    __shared__ float sGrad[kWaves][BATCH][9];
    __shared__ float sCk[kMaxCheckpoints][kThreads];

    if (ctl->overflow) return;
    const int tile = blockIdx.x;
    const int tileX = tile % tilesX, tileY = tile / tilesX;
    const int tid = threadIdx.x;
    const int wave = tid / kWave, lane = __lane_id();
    const int bx = tileX * TW + (wave % kWavesX) * 8, by = tileY * TH + (wave / kWavesX) * 8;
    const int x = bx + (lane & 7), y = by + (lane >> 3);
    const bool inside = x < width && y < height;
    const float px = x + 0.5f, py = y + 0.5f;
    const uint64_t below = lanesBelow();

    const uint2 range = ranges[tile];
    const uint32_t len = range.y - range.x;
    const int nBatches = static_cast<int>((len + BATCH - 1) / BATCH);
    const int ckEvery = max(1, (nBatches + kMaxCheckpoints - 1) / kMaxCheckpoints);

    // Stages batch b: records, Gaussian ids and the per-wavefront culling mask; then this
    // wavefront's list of splats to visit. Returns the number of splats in the batch.
    auto stage = [&](int b, int& count) -> int {
        const uint32_t start = range.x + static_cast<uint32_t>(b) * BATCH;
        const int cnt = min(static_cast<uint32_t>(BATCH), range.y - start);
        for (int k = tid; k < cnt; k += kThreads) {
            const uint32_t g = sortedGauss[start + k];
            const float4 a = rec0[g];
            const float4 bb = rec1[g];
            s0[k] = a;
            s1[k] = bb;
            s2[k] = rec2[g];
            sGauss[k] = g;
            uint32_t m = 0;
            if (bb.y >= 1.0f / 255.0f) {
                const float kq = fmaxf(2.0f * __logf(255.0f * bb.y), 0.0f) * 1.05f + 0.1f;
                const float rB_C = -a.w / bb.x, rB_A = -a.w / a.z;
#pragma unroll
                for (int w = 0; w < kWaves; ++w) {
                    const float x0 = tileX * TW + (w % kWavesX) * 8 + 0.5f;
                    const float y0 = tileY * TH + (w / kWavesX) * 8 + 0.5f;
                    const float dx0 = x0 - a.x, dy0 = y0 - a.y;
                    if (quadMinOverBox(a.z, a.w, bb.x, rB_C, rB_A, dx0, dx0 + 7.0f, dy0, dy0 + 7.0f) <= kq)
                        m |= 1u << w;
                }
            }
            sMask[k] = static_cast<uint8_t>(m);
        }
        __syncthreads();
        count = 0;
        for (int g0 = 0; g0 < cnt; g0 += kWave) {
            const int j = g0 + lane;
            const bool hit = j < cnt && ((sMask[j] >> wave) & 1u);
            const uint64_t bal = __ballot(hit);
            if (hit) sList[wave][count + __popcll(bal & below)] = static_cast<uint8_t>(j);
            count += __popcll(bal);
        }
        waveSync();
        return cnt;
    };

    // Pass 1: forward walk.
    float T = 1.0f;
    bool done = !inside;
    uint32_t lastContrib = 0;  // list position + 1 of this pixel's last contributing splat
    int nbDone = 0;
    for (int b = 0; b < nBatches; ++b) {
        if (__syncthreads_count(done) == kThreads) break;
        if (b % ckEvery == 0) sCk[b / ckEvery][tid] = T;
        int count;
        stage(b, count);
        nbDone = b + 1;
        if (__ballot(!done) == 0) continue;
        for (int i = 0; i < count; ++i) {
            const int j = sList[wave][i];
            const float4 a = s0[j];
            const float4 bb = s1[j];
            if (!done) {
                const float dx = px - a.x, dy = py - a.y;
                const float power = -0.5f * (a.z * dx * dx + bb.x * dy * dy) - a.w * dx * dy;
                if (!(power > 0.0f)) {
                    const float alpha = fminf(0.99f, bb.y * __expf(power));
                    if (alpha >= 1.0f / 255.0f) {
                        const float nextT = T * (1.0f - alpha);
                        if (nextT < 1e-4f) {
                            done = true;
                        } else {
                            T = nextT;
                            lastContrib = static_cast<uint32_t>(b) * BATCH + j + 1;
                        }
                    }
                }
            }
            if ((i & 7) == 7 && __ballot(!done) == 0) break;
        }
    }

    // Pass 2: reverse walk.
    float gr = 0.0f, gg = 0.0f, gb = 0.0f;
    if (inside) {
        const size_t p = static_cast<size_t>(y) * width + x;
        gr = gradImage[3 * p + 0];
        gg = gradImage[3 * p + 1];
        gb = gradImage[3 * p + 2];
    }
    float ar = 0.0f, ag = 0.0f, ab = 0.0f;  // accum
    float Tcur = T;
    for (int b = nbDone - 1; b >= 0; --b) {
        __syncthreads();
        for (int k = tid; k < kWaves * BATCH * 9; k += kThreads) (&sGrad[0][0][0])[k] = 0.0f;
        int count;
        const int cnt = stage(b, count);
        if (__ballot(lastContrib > static_cast<uint32_t>(b) * BATCH) != 0) {
            for (int i = count - 1; i >= 0; --i) {
                const int j = sList[wave][i];
                const float4 a = s0[j];
                const float4 bb = s1[j];
                float d[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
                bool contrib = false;
                if (static_cast<uint32_t>(b) * BATCH + j < lastContrib) {
                    const float dx = px - a.x, dy = py - a.y;
                    const float power = -0.5f * (a.z * dx * dx + bb.x * dy * dy) - a.w * dx * dy;
                    if (!(power > 0.0f)) {
                        const float G = __expf(power);
                        const float raw = bb.y * G;
                        const float alpha = fminf(0.99f, raw);
                        if (alpha >= 1.0f / 255.0f) {
                            contrib = true;
                            const float4 c = s2[j];
                            const float Ti = Tcur / (1.0f - alpha);
                            const float w = alpha * Ti;
                            d[6] = gr * w;
                            d[7] = gg * w;
                            d[8] = gb * w;
                            const float gAlpha = Ti * (gr * (c.x - ar) + gg * (c.y - ag) + gb * (c.z - ab));
                            ar = alpha * c.x + (1.0f - alpha) * ar;
                            ag = alpha * c.y + (1.0f - alpha) * ag;
                            ab = alpha * c.z + (1.0f - alpha) * ab;
                            if (raw <= 0.99f) {
                                d[5] = gAlpha * G;
                                const float gPow = gAlpha * raw;
                                d[0] = gPow * (a.z * dx + a.w * dy);
                                d[1] = gPow * (bb.x * dy + a.w * dx);
                                d[2] = gPow * (-0.5f * dx * dx);
                                d[3] = gPow * (-dx * dy);
                                d[4] = gPow * (-0.5f * dy * dy);
                            }
                            Tcur = Ti;
                        }
                    }
                }
                if (__ballot(contrib) != 0) {
#pragma unroll
                    for (int c = 0; c < 9; ++c) {
                        const float s = waveSumLast(d[c]);
                        if (lane == kWave - 1) sGrad[wave][j][c] = s;
                    }
                }
            }
        }
        if (b % ckEvery == 0) Tcur = sCk[b / ckEvery][tid];
        __syncthreads();
        // One partial per pair of this batch: the wavefronts' sums in fixed order.
        for (int k = tid; k < cnt; k += kThreads) {
            const uint32_t g = sGauss[k];
            const uint2 rp = rect[g];
            const uint32_t x0 = rp.x & 0xffffu, y0 = rp.x >> 16, w = (rp.y & 0xffffu) - x0 + 1;
            const uint32_t slot = pairStart[g] + (static_cast<uint32_t>(tileY) - y0) * w + (static_cast<uint32_t>(tileX) - x0);
            float* out = partial + static_cast<size_t>(slot) * 9;
#pragma unroll
            for (int c = 0; c < 9; ++c) {
                float s = 0.0f;
#pragma unroll
                for (int wv = 0; wv < kWaves; ++wv) s += sGrad[wv][k][c];
                out[c] = s;
            }
        }
    }
    // Pairs after the point where every pixel had stopped get zero partials.
    for (uint32_t k = static_cast<uint32_t>(nbDone) * BATCH + tid; k < len; k += kThreads) {
        const uint32_t g = sortedGauss[range.x + k];
        const uint2 rp = rect[g];
        const uint32_t x0 = rp.x & 0xffffu, y0 = rp.x >> 16, w = (rp.y & 0xffffu) - x0 + 1;
        const uint32_t slot = pairStart[g] + (static_cast<uint32_t>(tileY) - y0) * w + (static_cast<uint32_t>(tileX) - x0);
        float* out = partial + static_cast<size_t>(slot) * 9;
#pragma unroll
        for (int c = 0; c < 9; ++c) out[c] = 0.0f;
    }
}

// d basis_k / d(x, y, z) for the basis of shBasis (product rule on each polynomial).
__host__ __device__ inline void shBasisGrad(int degree, float x, float y, float z, float g[16][3]) {
#pragma unroll
    for (int k = 0; k < 16; ++k) g[k][0] = g[k][1] = g[k][2] = 0.0f;
    if (degree >= 1) {
        g[1][1] = -kShC1;
        g[2][2] = kShC1;
        g[3][0] = -kShC1;
    }
    if (degree >= 2) {
        g[4][0] = kShC2[0] * y; g[4][1] = kShC2[0] * x;
        g[5][1] = kShC2[1] * z; g[5][2] = kShC2[1] * y;
        g[6][0] = kShC2[2] * (-2 * x); g[6][1] = kShC2[2] * (-2 * y); g[6][2] = kShC2[2] * (4 * z);
        g[7][0] = kShC2[3] * z; g[7][2] = kShC2[3] * x;
        g[8][0] = kShC2[4] * (2 * x); g[8][1] = kShC2[4] * (-2 * y);
        if (degree >= 3) {
            const float xx = x * x, yy = y * y, zz = z * z;
            g[9][0] = kShC3[0] * y * 6 * x; g[9][1] = kShC3[0] * (3 * xx - 3 * yy);
            g[10][0] = kShC3[1] * y * z; g[10][1] = kShC3[1] * x * z; g[10][2] = kShC3[1] * x * y;
            g[11][0] = kShC3[2] * y * (-2 * x); g[11][1] = kShC3[2] * (4 * zz - xx - 3 * yy); g[11][2] = kShC3[2] * y * 8 * z;
            g[12][0] = kShC3[3] * z * (-6 * x); g[12][1] = kShC3[3] * z * (-6 * y);
            g[12][2] = kShC3[3] * (6 * zz - 3 * xx - 3 * yy);
            g[13][0] = kShC3[4] * (4 * zz - 3 * xx - yy); g[13][1] = kShC3[4] * x * (-2 * y); g[13][2] = kShC3[4] * x * 8 * z;
            g[14][0] = kShC3[5] * z * 2 * x; g[14][1] = kShC3[5] * z * (-2 * y); g[14][2] = kShC3[5] * (xx - yy);
            g[15][0] = kShC3[6] * (3 * xx - 3 * yy); g[15][1] = kShC3[6] * x * (-6 * y);
        }
    }
}

// Adjoint of projectGaussian (raw parameters): from gs = dL/d{u, v, A, B, C, opacity, r, g, b}
// to the gradients of the mean (gm), log-scales (gls), raw quaternion (gq), opacity logit and the
// SH coefficients (gsh, [coefficient][channel], zero past the evaluated degree). shi holds the
// Gaussian's SH, [coefficient][channel].
__host__ __device__ inline void gaussAdjoint(const GaussProj& g, const CamParams& cam, const float gs[9],
                                             const float* shi, float gm[3], float gls[3], float gq[4], float* glogit,
                                             float* gsh) {
    const float gu = gs[0], gv = gs[1], gA = gs[2], gB = gs[3], gC = gs[4], go = gs[5];
    const float px = g.px, py = g.py, pz = g.pz;

    float gp0 = gu * cam.fx / pz, gp1 = gv * cam.fy / pz;
    float gp2 = -gu * cam.fx * px / (pz * pz) - gv * cam.fy * py / (pz * pz);
    // Conic A = c / det, B = -b / det, C = a / det of the 2D covariance (det = a c - b^2).
    const float idet = 1.0f / g.det, idet2 = idet * idet;
    const float ga = gA * (-g.c * g.c * idet2) + gB * (g.b * g.c * idet2) + gC * (-g.b * g.b * idet2);
    const float gbb = gA * (2.0f * g.b * g.c * idet2) + gB * (-idet - 2.0f * g.b * g.b * idet2) + gC * (2.0f * g.a * g.b * idet2);
    const float gc = gA * (-g.b * g.b * idet2) + gB * (g.a * g.b * idet2) + gC * (-g.a * g.a * idet2);
    // cov = T Sigma T^T with a = cov00, b = cov01 = cov10, c = cov11.
    const float G[2][2] = {{ga, 0.5f * gbb}, {0.5f * gbb, gc}};
    float TS[2][3];
    for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 3; ++c) TS[r][c] = g.T[r][0] * g.Sigma[0][c] + g.T[r][1] * g.Sigma[1][c] + g.T[r][2] * g.Sigma[2][c];
    float GT[2][3];  // G T
    for (int r = 0; r < 2; ++r)
        for (int j = 0; j < 3; ++j) GT[r][j] = G[r][0] * g.T[0][j] + G[r][1] * g.T[1][j];
    float gSigma[3][3];
    for (int i2 = 0; i2 < 3; ++i2)
        for (int j = 0; j < 3; ++j) gSigma[i2][j] = g.T[0][i2] * GT[0][j] + g.T[1][i2] * GT[1][j];
    float gT[2][3];
    for (int r = 0; r < 2; ++r)
        for (int j = 0; j < 3; ++j) gT[r][j] = 2.0f * (G[r][0] * TS[0][j] + G[r][1] * TS[1][j]);
    // T = J W: dJ = dT W^T.
    float gJ[2][3];
    for (int r = 0; r < 2; ++r)
        for (int k = 0; k < 3; ++k) gJ[r][k] = gT[r][0] * cam.R[3 * k] + gT[r][1] * cam.R[3 * k + 1] + gT[r][2] * cam.R[3 * k + 2];
    const float pz2 = pz * pz, pz3 = pz2 * pz;
    gp2 += gJ[0][0] * (-cam.fx / pz2) + gJ[1][1] * (-cam.fy / pz2) + gJ[0][2] * (2.0f * cam.fx * g.tx / pz3) +
           gJ[1][2] * (2.0f * cam.fy * g.ty / pz3);
    const float gtx = gJ[0][2] * (-cam.fx / pz2), gty = gJ[1][2] * (-cam.fy / pz2);
    if (!g.clampX) gp0 += gtx; else gp2 += gtx * (g.tx / pz);
    if (!g.clampY) gp1 += gty; else gp2 += gty * (g.ty / pz);
    for (int j = 0; j < 3; ++j) gm[j] = cam.R[j] * gp0 + cam.R[3 + j] * gp1 + cam.R[6 + j] * gp2;
    // Sigma = M M^T, M = Rq diag(s), s = exp(log-scale) * splatScale.
    float gRq[3][3], gsc[3] = {0, 0, 0};
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            const float gM = 2.0f * (gSigma[r][0] * g.M[0][c] + gSigma[r][1] * g.M[1][c] + gSigma[r][2] * g.M[2][c]);
            gRq[r][c] = gM * g.s[c];
            gsc[c] += gM * g.Rq[r][c];
        }
    for (int j = 0; j < 3; ++j) gls[j] = gsc[j] * g.s[j];
    const float w = g.q[0], x = g.q[1], y = g.q[2], z = g.q[3];
    float gn[4];
    gn[0] = 2.0f * (-z * gRq[0][1] + y * gRq[0][2] + z * gRq[1][0] - x * gRq[1][2] - y * gRq[2][0] + x * gRq[2][1]);
    gn[1] = 2.0f * (y * gRq[0][1] + z * gRq[0][2] + y * gRq[1][0] - 2.0f * x * gRq[1][1] - w * gRq[1][2] +
                    z * gRq[2][0] + w * gRq[2][1] - 2.0f * x * gRq[2][2]);
    gn[2] = 2.0f * (-2.0f * y * gRq[0][0] + x * gRq[0][1] + w * gRq[0][2] + x * gRq[1][0] + z * gRq[1][2] -
                    w * gRq[2][0] + z * gRq[2][1] - 2.0f * y * gRq[2][2]);
    gn[3] = 2.0f * (-2.0f * z * gRq[0][0] - w * gRq[0][1] + x * gRq[0][2] + w * gRq[1][0] - 2.0f * z * gRq[1][1] +
                    y * gRq[1][2] + x * gRq[2][0] + y * gRq[2][1]);
    // q = q_raw / |q_raw|.
    if (g.qlen > 0.0f) {
        const float dot = gn[0] * w + gn[1] * x + gn[2] * y + gn[3] * z;
        for (int j = 0; j < 4; ++j) gq[j] = (gn[j] - dot * g.q[j]) / g.qlen;
    } else {
        for (int j = 0; j < 4; ++j) gq[j] = 0.0f;
    }
    // opacity = sigmoid(logit).
    *glogit = go * g.opacity * (1.0f - g.opacity);
    // SH colour, clamped at zero: d sh_k = basis_k d rgb; d dir through the basis.
    const float grgb[3] = {g.rgbRaw.x > 0.0f ? gs[6] : 0.0f, g.rgbRaw.y > 0.0f ? gs[7] : 0.0f,
                           g.rgbRaw.z > 0.0f ? gs[8] : 0.0f};
    float bgrad[16][3];
    shBasisGrad(cam.shDegree, g.dir[0], g.dir[1], g.dir[2], bgrad);
    float gdir[3] = {0, 0, 0};
    for (int k = 0; k < 16; ++k) {
        if (k < g.shCount) {
            float gbk = 0.0f;
            for (int c = 0; c < 3; ++c) {
                gsh[k * 3 + c] = grgb[c] * g.basis[k];
                gbk += grgb[c] * shi[k * 3 + c];
            }
            for (int j = 0; j < 3; ++j) gdir[j] += gbk * bgrad[k][j];
        }
    }
    for (int j = g.shCount * 3; j < cam.shCoeffs * 3; ++j) gsh[j] = 0.0f;
    // dir = (m - eye) / |m - eye|.
    const float dd = gdir[0] * g.dir[0] + gdir[1] * g.dir[1] + gdir[2] * g.dir[2];
    for (int j = 0; j < 3; ++j) gm[j] += (gdir[j] - dd * g.dir[j]) / g.dlen;
}

// Per Gaussian: sums its pair partials in slot order, then runs gaussAdjoint to the gradients of
// means, log-scales, quaternions (w, x, y, z), opacity logits and SH [coefficient][channel].
// Culled Gaussians get zero gradients.
static __global__ __launch_bounds__(256) void k_gaussBackward(int n, const float* __restrict__ pos,
                                                              const float* __restrict__ logScale,
                                                              const float4* __restrict__ quat,
                                                              const float* __restrict__ logit,
                                                              const float* __restrict__ sh, CamParams cam,
                                                              const uint32_t* __restrict__ tilesTouched,
                                                              const uint32_t* __restrict__ pairStart,
                                                              const float* __restrict__ partial,
                                                              float* __restrict__ gMeans, float* __restrict__ gScales,
                                                              float* __restrict__ gQuats, float* __restrict__ gLogit,
                                                              float* __restrict__ gSh,
                                                              const FrameCtl* __restrict__ ctl) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n || ctl->overflow) return;
    const int K3 = cam.shCoeffs * 3;
    float* gshi = gSh + static_cast<size_t>(i) * K3;
    const uint32_t tt = tilesTouched[i];
    GaussProj g;
    if (tt == 0 || !projectGaussian<false, true>(i, n, pos, logScale, quat, logit, sh, cam, g)) {
        for (int j = 0; j < 3; ++j) gMeans[3 * i + j] = 0.0f, gScales[3 * i + j] = 0.0f;
        for (int j = 0; j < 4; ++j) gQuats[4 * i + j] = 0.0f;
        gLogit[i] = 0.0f;
        for (int j = 0; j < K3; ++j) gshi[j] = 0.0f;
        return;
    }
    float gs[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const float* src = partial + static_cast<size_t>(pairStart[i]) * 9;
    for (uint32_t k = 0; k < tt; ++k)
#pragma unroll
        for (int c = 0; c < 9; ++c) gs[c] += src[9 * k + c];
    float gm[3], gls[3], gq[4], gl;
    gaussAdjoint(g, cam, gs, sh + static_cast<size_t>(i) * K3, gm, gls, gq, &gl, gshi);
    for (int j = 0; j < 3; ++j) gMeans[3 * i + j] = gm[j], gScales[3 * i + j] = gls[j];
    for (int j = 0; j < 4; ++j) gQuats[4 * i + j] = gq[j];
    gLogit[i] = gl;
}

}  // namespace ans
