// Backward kernels for gfx950: the adjoint of the tile blend and of the per-Gaussian projection.
// There is no backward in the reference (its training uses a third-party rasterizer); the
// forward these differentiate is the reference's (src/render.cpp:39-104, :193-221).
//
// Gradients are reduced without atomics: every (tile, Gaussian) pair is owned by one workgroup,
// which sums its pixels' contributions in a fixed order and writes one partial record per pair
// into the pair's slot; slots are numbered in Gaussian index order, each Gaussian's tiles
// row-major ([slotStart, slotStart + tiles touched)), so the per-Gaussian kernel reads one
// contiguous run per workgroup and sums each Gaussian's records in order.
#pragma once

#include "forward.hpp"

namespace ans {

// Parameter and gradient arrays of a differentiable call, read by the kernels from the
// per-call device table (so that a captured graph replays for any tensors): the layout of
// ans_params / ans_grads (means [n][3], log-scales [n][3], quaternions [n][4], opacity logits [n],
// SH [n][coefficient][channel]). A batched call reads it from its per-call device table.
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


template <int CTRL>
__device__ inline float dppAdd(float v) {
    return v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false));
}

__device__ inline float addSwapped(unsigned a, unsigned b) { return __uint_as_float(a) + __uint_as_float(b); }

// Sum over the 64 lanes, in every lane: a butterfly of full-mask DPP adds within each 16-lane
// row (quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror, row_mirror), then gfx950's
// permlane16/permlane32 swaps across rows. Lane 63 is read by the callers; the order of the
// additions is fixed, so the result is deterministic.
__device__ inline float waveSum(float v) {
    v = dppAdd<0xB1>(v);
    v = dppAdd<0x4E>(v);
    v = dppAdd<0x141>(v);
    v = dppAdd<0x140>(v);
    const auto r16 = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    v = addSwapped(r16[0], r16[1]);
    const auto r32 = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    return addSwapped(r32[0], r32[1]);
}

// v = sel ? a : b in this lane, w = sel ? b : a, and v plus w from the DPP partner lane: where
// sel is true in a lane and false in its partner (and the other way round), each lane ends up with
// a pairwise sum of a or of b.
template <int CTRL>
__device__ inline float foldPair(float a, float b, bool sel) {
    const float v = sel ? a : b, w = sel ? b : a;
    return v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(w), CTRL, 0xf, 0xf, false));
}

// Sums over the 64 lanes of two splats' 9 gradient components at once, a reduce-scatter: each
// stage halves the lanes a value is spread over and the number of registers holding values.
// permlane32_swap leaves lanes 0-31 holding splat 1's pairwise sums and lanes 32-63 splat 2's;
// permlane16_swap pairs components (k, k + 5), k < 4, so that rows 0..3 of register k hold
// (splat 1, k), (splat 1, k + 5), (splat 2, k), (splat 2, k + 5) (component 4 pairs with itself:
// rows 0, 1 and 2, 3 repeat it); within each 16-lane row, row_ror:8 folds registers (0, 1) and
// (2, 3), row_half_mirror the two results, quad_perm [2,3,0,1] that with component 4, and
// quad_perm [1,0,3,2] completes every sum. Lane p of row r then holds, for p % 4 == 0, component
// {0, 2, 1, 3}[p / 4] (+ 5 in rows 1 and 3), and for p % 4 == 2 component 4, of splat 1 (rows 0, 1)
// or splat 2 (rows 2, 3); the lanes with p % 4 == 0, and p == 2 in rows 0 and 2, store their
// sum at rowBase1 or rowBase2 + component. The order of additions is fixed.
__device__ inline void pairSumStore(const float (&d1)[9], const float (&d2)[9], float* rowBase1, float* rowBase2) {
    float t[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) {
        const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(d1[c]), __float_as_uint(d2[c]), false, false);
        t[c] = addSwapped(r[0], r[1]);
    }
    float u[5];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(t[k]), __float_as_uint(t[k + 5]), false, false);
        u[k] = addSwapped(r[0], r[1]);
    }
    {
        const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(t[4]), __float_as_uint(t[4]), false, false);
        u[4] = addSwapped(r[0], r[1]);
    }
    const int lane = __lane_id(), p = lane & 15;
    const float v01 = foldPair<0x128>(u[0], u[1], p < 8);  // row_ror:8
    const float v23 = foldPair<0x128>(u[2], u[3], p < 8);
    const float v4 = dppAdd<0x128>(u[4]);
    const float w = foldPair<0x141>(v01, v23, (p & 7) < 4);  // row_half_mirror
    const float w4 = dppAdd<0x141>(v4);
    const float z = dppAdd<0xB1>(foldPair<0x4E>(w, w4, (p & 3) < 2));  // quad_perm [2,3,0,1], [1,0,3,2]
    // An empty use of the sum outside the branch below: otherwise the compiler moves the last DPP
    // add into the branch and splits it into a register clear, a DPP move and an add.
    asm volatile("" : : "v"(z));
    const int row = lane >> 4;
    const bool writer = (p & 3) == 0 || (p == 2 && (row & 1) == 0);
    if (writer) {
        const int k = (p & 3) ? 4 : ((p >> 2) == 1 ? 2 : (p >> 2) == 2 ? 1 : (p >> 2)) + (row & 1) * 5;
        (row < 2 ? rowBase1 : rowBase2)[k] = z;
    }
}

constexpr int kPartialStride = 12;  // floats per pair record: u v A B C opacity r g b, 3 pad

// Adjoint of k_blend for one tile, one lane per pixel as in the forward. The forward left, per
// pixel, its final transmittance T and the list position of the last splat it blended; the walk
// goes back to front over the batches up to the tile's furthest such position, recovering each
// splat's transmittance as T_i = T_{i+1} / (1 - alpha_i) from the same alpha the forward
// computed. Per contributing splat i of a pixel with image gradient g = dL/dC:
//   dL/drgb_i = g alpha_i T_i,  dL/dalpha_i = T_i g . (rgb_i - accum),  accum <- alpha_i rgb_i +
//   (1 - alpha_i) accum  (accum = colour behind i per unit of the transmittance after i);
//   alpha = o G (no gradient where the 0.99 clamp is active), G = exp(power),
//   power = -(A dx^2 + C dy^2) / 2 - B dx dy, dx = px - u, dy = py - v.
// Each wavefront visits only the splats whose culling test (the forward's) admits its 8 x 8
// block, two at a time: pairSumStore reduces the two splats' per-pixel moments over the
// wavefront, and the wavefronts' sums are added in fixed order and turned into one partial
// record per pair: partial[slot * 12 + {u, v, A, B, C, opacity, r, g, b}].
template <int TW, int TH, int BATCH>
__device__ inline void blendBackwardTile(int tileX, int tileY, uint2 range, const uint32_t* __restrict__ sortedGauss,
                                         const SplatRec* __restrict__ rec, const uint32_t* __restrict__ slotStart,
                                         const uint2* __restrict__ pixState, int width, int height,
                                         const float* __restrict__ gradImage, float* __restrict__ partial) {
    constexpr int kThreads = TW * TH;
    constexpr int kWaves = kThreads / kWave;
    constexpr int kWavesX = TW / 8;
    // Staged per splat in three arrays of 16-byte rows read at one byte offset (the visit lists
    // hold offsets, so the walk computes no addresses): (u, v, qa, qb), (qc, log2 o, r, g),
    // (b, 1 / o). Row BATCH is a splat that never blends (log2 o = -inf, zero colour): the
    // partner of the last splat of an odd-length list, which every list has at its entry 0.
    __shared__ float4 sRec[3][BATCH + 1];
    __shared__ uint8_t sMask[BATCH];
    __shared__ uint16_t sList[kWaves][BATCH + 1];
    __shared__ uint32_t sSlot[BATCH];
    __shared__ float sGrad[kWaves][BATCH + 1][9];  // row BATCH: the never-blending splat's zeros
    __shared__ uint32_t sMaxLast;
    auto row = [&](int r, uint32_t off) -> const float4& {
        return *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(sRec[r]) + off);
    };
    constexpr uint16_t kNone = BATCH * 16;

    const int tid = threadIdx.x;
    const int wave = tid / kWave, lane = __lane_id();
    const int bx = tileX * TW + (wave % kWavesX) * 8, by = tileY * TH + (wave / kWavesX) * 8;
    const int x = bx + (lane & 7), y = by + (lane >> 3);
    const bool inside = x < width && y < height;
    const float px = x + 0.5f, py = y + 0.5f;
    const uint64_t below = lanesBelow();

    const uint32_t len = range.y - range.x;

    // Stages batch b: records, each pair's partial slot and the per-wavefront culling mask; then
    // this wavefront's list of splats to visit. Returns the number of splats in the batch.
    auto stage = [&](int b, int& count) -> int {
        const uint32_t start = range.x + static_cast<uint32_t>(b) * BATCH;
        const int cnt = min(static_cast<uint32_t>(BATCH), range.y - start);
        for (int k = tid; k < cnt; k += kThreads) {
            const uint32_t g = sortedGauss[start + k];
            const float4 a = rec[g].geo;
            const float4 bb = rec[g].geo2;
            const uint4 meta = rec[g].meta;
            const uint2 rp = make_uint2(meta.x, meta.y);
            const uint32_t first = slotStart[g];
            const float3 q = stagedConic(a.z, a.w, bb.x);  // the blend's staged conic and log2 o
            const float4 col = rec[g].col;
            sRec[0][k] = make_float4(a.x, a.y, q.x, q.y);
            sRec[1][k] = make_float4(q.z, stagedLogOpacity(bb.y), col.x, col.y);
            sRec[2][k] = make_float4(col.z, 1.0f / bb.y, 0.0f, 0.0f);
            const uint32_t x0 = rp.x & 0xffffu, y0 = rp.x >> 16, w = (rp.y & 0xffffu) - x0 + 1;
            sSlot[k] = first + (static_cast<uint32_t>(tileY) - y0) * w + (static_cast<uint32_t>(tileX) - x0);
            uint32_t m = 0;
            if (bb.y >= 1.0f / 255.0f) {
                const float kq = fmaxf(2.0f * __logf(255.0f * bb.y), 0.0f) * 1.05f + 0.1f;
                const float rB_C = -a.w / bb.x, rB_A = -a.w / a.z;
#pragma unroll
                for (int wv = 0; wv < kWaves; ++wv) {
                    const float x0f = tileX * TW + (wv % kWavesX) * 8 + 0.5f;
                    const float y0f = tileY * TH + (wv / kWavesX) * 8 + 0.5f;
                    const float dx0 = x0f - a.x, dy0 = y0f - a.y;
                    if (quadMinOverBox(a.z, a.w, bb.x, rB_C, rB_A, dx0, dx0 + 7.0f, dy0, dy0 + 7.0f) <= kq)
                        m |= 1u << wv;
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
            if (hit) sList[wave][1 + count + __popcll(bal & below)] = static_cast<uint16_t>(j * 16);
            count += __popcll(bal);
        }
        waveSync();
        return cnt;
    };

    // The forward's per-pixel end state, and the furthest last-blended position in the tile.
    float T = 1.0f;
    uint32_t lastContrib = 0;  // list position + 1 of this pixel's last blended splat
    if (inside) {
        const uint2 st = pixState[static_cast<size_t>(y) * width + x];
        T = __uint_as_float(st.x);
        lastContrib = st.y;
    }
    if (tid == 0) {
        sMaxLast = 0;
        sRec[0][BATCH] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        sRec[1][BATCH] = make_float4(0.0f, -__builtin_huge_valf(), 0.0f, 0.0f);
        sRec[2][BATCH] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    if (lane == 0) sList[wave][0] = kNone;
    uint32_t m = lastContrib;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) m = max(m, static_cast<uint32_t>(__shfl_xor(static_cast<int>(m), d, kWave)));
    __syncthreads();
    if (lane == 0) atomicMax(&sMaxLast, m);
    __syncthreads();
    const int nbDone = static_cast<int>((sMaxLast + BATCH - 1) / BATCH);

    // Pass 2: reverse walk.
    float gr = 0.0f, gg = 0.0f, gb = 0.0f;
    if (inside) {
        const size_t p = static_cast<size_t>(y) * width + x;
        gr = gradImage[3 * p + 0];
        gg = gradImage[3 * p + 1];
        gb = gradImage[3 * p + 2];
    }
    float S = 0.0f;  // g . accum
    int lastLocal = 0;  // staged byte offsets below it are positions this pixel blended
    float Tcur = T;
    // This pixel's contributions from the splat staged at byte offset off (position off / 16 of
    // the batch; zero if it does not blend there), as the moments the pair record is formed from:
    // d = (gPow dx, gPow dy, gPow dx dx, gPow dx dy, gPow dy dy, gPow, dL/drgb) with
    // gPow = dL/dpower; steps T and g . accum back past the splat. Returns alpha if it
    // contributed, else 0.
    // power and alpha are the blend's, bit for bit (same staged conic, same operations), so every
    // blend/skip decision matches the forward's; selects and zero weights replace branches.
    auto visit = [&](uint32_t off, float (&d)[9]) -> float {
        const float4 a = row(0, off), bb = row(1, off), c = row(2, off);
        const float dx = px - a.x, dy = py - a.y;
        const float e = splatExponent(a.z, a.w, bb.x, bb.y, dx, dy);
        const float raw = __builtin_amdgcn_exp2f(e);  // o G
        const float alpha = fminf(0.99f, raw);
        const bool use = static_cast<int>(off) < lastLocal && !(e > bb.y) && alpha >= 1.0f / 255.0f;
        const float Ti = Tcur * __builtin_amdgcn_rcpf(1.0f - alpha);
        // A zero weight leaves every sum below unchanged exactly (all terms are finite).
        const float wa = use ? alpha : 0.0f;
        const float w = wa * Ti;
        d[6] = gr * w;
        d[7] = gg * w;
        d[8] = gb * w;
        // dL/dalpha = T_i g . (rgb - accum), with g . accum carried as one running sum:
        // accum <- accum + alpha (rgb - accum) gives g . accum <- S + alpha (g . rgb - S).
        const float diff = __builtin_fmaf(gb, c.x, __builtin_fmaf(gg, bb.w, gr * bb.z)) - S;
        const float gAlpha = Ti * diff;
        // No gradient where the 0.99 clamp is active.
        const bool open = use && raw <= 0.99f;
        const float gPow = open ? gAlpha * raw : 0.0f;
        const float p = gPow * dx, q = gPow * dy;
        d[0] = p;
        d[1] = q;
        d[2] = p * dx;
        d[3] = p * dy;
        d[4] = q * dy;
        d[5] = gPow;
        S = __builtin_fmaf(wa, diff, S);
        Tcur = use ? Ti : Tcur;
        return wa;
    };
    for (int b = nbDone - 1; b >= 0; --b) {
        __syncthreads();
        for (int k = tid; k < kWaves * (BATCH + 1) * 9; k += kThreads) (&sGrad[0][0][0])[k] = 0.0f;
        int count;
        const int cnt = stage(b, count);
        if (__ballot(lastContrib > static_cast<uint32_t>(b) * BATCH) != 0 && count > 0) {
            // Pairs (i, i - 1) back to front over list entries count..1 (entry 0: the splat that
            // never blends). Records are read at the top of each step: carried across steps they
            // cost more register moves than the LDS latency other waves hide.
            lastLocal = (static_cast<int>(lastContrib) - b * BATCH) * 16;
            for (int i = count; i >= 1; i -= 2) {
                const uint32_t o1 = sList[wave][i], o2 = sList[wave][i - 1];
                float d1[9], d2[9];
                const float w1 = visit(o1, d1);
                const float w2 = visit(o2, d2);
                // A contributing splat has alpha >= 1/255: the ballot of a fresh compare costs
                // one instruction, that of the carried flag three.
                if (__builtin_amdgcn_ballot_w64(w1 > 0.0f || w2 > 0.0f) != 0)
                    pairSumStore(d1, d2, &sGrad[wave][o1 >> 4][0], &sGrad[wave][o2 >> 4][0]);
            }
        }
        __syncthreads();
        // One partial per pair of this batch: the wavefronts' moment sums in fixed order, turned
        // into the record's gradients. With the staged qa = -log2(e) A/2, qb = -log2(e) B,
        // qc = -log2(e) C/2 and power = -(A dx^2 + C dy^2) / 2 - B dx dy:
        //   dL/du = sum gPow (A dx + B dy) = -(2 qa M_x + qb M_y) ln 2,
        //   dL/dv = sum gPow (B dx + C dy) = -(qb M_x + 2 qc M_y) ln 2,
        //   dL/dA = -M_xx / 2, dL/dB = -M_xy, dL/dC = -M_yy / 2, dL/do = M_1 / o,
        // M_x = sum gPow dx and so on.
        for (int k = tid; k < cnt; k += kThreads) {
            float s[9];
#pragma unroll
            for (int c = 0; c < 9; ++c) {
                s[c] = 0.0f;
#pragma unroll
                for (int wv = 0; wv < kWaves; ++wv) s[c] += sGrad[wv][k][c];
            }
            const float4 q = sRec[0][k];
            const float qc = sRec[1][k].x, invO = sRec[2][k].y;
            constexpr float kLn2 = 1.0f / kLog2e;
            const float gu = -kLn2 * __builtin_fmaf(q.z + q.z, s[0], q.w * s[1]);
            const float gv = -kLn2 * __builtin_fmaf(q.w, s[0], (qc + qc) * s[1]);
            float4* out = reinterpret_cast<float4*>(partial + static_cast<size_t>(sSlot[k]) * kPartialStride);
            out[0] = make_float4(gu, gv, -0.5f * s[2], -s[3]);
            out[1] = make_float4(-0.5f * s[4], invO * s[5], s[6], s[7]);
            out[2] = make_float4(s[8], 0.0f, 0.0f, 0.0f);
        }
    }
    // Pairs after the point where every pixel had stopped get zero partials.
    for (uint32_t k = static_cast<uint32_t>(nbDone) * BATCH + tid; k < len; k += kThreads) {
        const uint32_t g = sortedGauss[range.x + k];
        const uint4 m = rec[g].meta;
        const uint2 rp = make_uint2(m.x, m.y);
        const uint32_t x0 = rp.x & 0xffffu, y0 = rp.x >> 16, w = (rp.y & 0xffffu) - x0 + 1;
        const uint32_t slot = slotStart[g] + (static_cast<uint32_t>(tileY) - y0) * w + (static_cast<uint32_t>(tileX) - x0);
        float4* out = reinterpret_cast<float4*>(partial + static_cast<size_t>(slot) * kPartialStride);
        out[0] = out[1] = out[2] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}

template <int TW, int TH, int BATCH>
__global__ __launch_bounds__(TW* TH) void k_blendBackward(
    const uint2* __restrict__ ranges, const uint32_t* __restrict__ sortedGauss, const SplatRec* __restrict__ rec,
    const uint32_t* __restrict__ slotStart, const uint2* __restrict__ pixState, int width, int height, int tilesX,
    const float* __restrict__ gradImage, float* __restrict__ partial, const FrameCtl* __restrict__ ctl) {
    if (ctl->overflow) return;
    const int tile = blockIdx.x;
    blendBackwardTile<TW, TH, BATCH>(tile % tilesX, tile / tilesX, ranges[tile], sortedGauss, rec, slotStart, pixState,
                                     width, height, gradImage, partial);
}

// Row k of the SH basis Jacobian: d basis_k / d (x, y, z) for the real SH of sh.cpp:18-58
// (rows past the degree the caller evaluates are never asked for). With k a constant in an
// unrolled loop the switch folds away, so the 16 x 3 Jacobian is never held at once.
__host__ __device__ inline void shBasisGradRow(int k, float x, float y, float z, float g[3]) {
    const float xx = x * x, yy = y * y, zz = z * z;
    g[0] = g[1] = g[2] = 0.0f;
    switch (k) {
        case 1: g[1] = -kShC1; break;
        case 2: g[2] = kShC1; break;
        case 3: g[0] = -kShC1; break;
        case 4: g[0] = kShC2[0] * y; g[1] = kShC2[0] * x; break;
        case 5: g[1] = kShC2[1] * z; g[2] = kShC2[1] * y; break;
        case 6: g[0] = kShC2[2] * (-2 * x); g[1] = kShC2[2] * (-2 * y); g[2] = kShC2[2] * (4 * z); break;
        case 7: g[0] = kShC2[3] * z; g[2] = kShC2[3] * x; break;
        case 8: g[0] = kShC2[4] * (2 * x); g[1] = kShC2[4] * (-2 * y); break;
        case 9: g[0] = kShC3[0] * y * 6 * x; g[1] = kShC3[0] * (3 * xx - 3 * yy); break;
        case 10: g[0] = kShC3[1] * y * z; g[1] = kShC3[1] * x * z; g[2] = kShC3[1] * x * y; break;
        case 11:
            g[0] = kShC3[2] * y * (-2 * x); g[1] = kShC3[2] * (4 * zz - xx - 3 * yy); g[2] = kShC3[2] * y * 8 * z;
            break;
        case 12:
            g[0] = kShC3[3] * z * (-6 * x); g[1] = kShC3[3] * z * (-6 * y); g[2] = kShC3[3] * (6 * zz - 3 * xx - 3 * yy);
            break;
        case 13:
            g[0] = kShC3[4] * (4 * zz - 3 * xx - yy); g[1] = kShC3[4] * x * (-2 * y); g[2] = kShC3[4] * x * 8 * z;
            break;
        case 14: g[0] = kShC3[5] * z * 2 * x; g[1] = kShC3[5] * z * (-2 * y); g[2] = kShC3[5] * (xx - yy); break;
        case 15: g[0] = kShC3[6] * (3 * xx - 3 * yy); g[1] = kShC3[6] * x * (-6 * y); break;
        default: break;
    }
}

// The view-dependent part of gaussAdjoint: the mean gradient gm (overwritten), dL/dSigma of the
// 3D covariance (gSigma, overwritten) and the SH gradients (gsh, stored or with ACC added).
template <bool ACC>
__host__ __device__ inline void gaussAdjointView(const GaussProj& g, const CamParams& cam, const float gs[9],
                                                 const float* shi, float gm[3], float gSigma[3][3], float* gsh) {
    const float gu = gs[0], gv = gs[1], gA = gs[2], gB = gs[3], gC = gs[4];
    const float px = g.px, py = g.py, pz = g.pz;
    // Powers of 1 / z are formed once and multiplied (one division instead of eleven).
    const float ipz = 1.0f / pz, ipz2 = ipz * ipz, ipz3 = ipz2 * ipz;

    float gp0 = gu * cam.fx * ipz, gp1 = gv * cam.fy * ipz;
    float gp2 = -gu * cam.fx * px * ipz2 - gv * cam.fy * py * ipz2;
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
    for (int i2 = 0; i2 < 3; ++i2)
        for (int j = 0; j < 3; ++j) gSigma[i2][j] = g.T[0][i2] * GT[0][j] + g.T[1][i2] * GT[1][j];
    float gT[2][3];
    for (int r = 0; r < 2; ++r)
        for (int j = 0; j < 3; ++j) gT[r][j] = 2.0f * (G[r][0] * TS[0][j] + G[r][1] * TS[1][j]);
    // T = J W: dJ = dT W^T.
    float gJ[2][3];
    for (int r = 0; r < 2; ++r)
        for (int k = 0; k < 3; ++k) gJ[r][k] = gT[r][0] * cam.R[3 * k] + gT[r][1] * cam.R[3 * k + 1] + gT[r][2] * cam.R[3 * k + 2];
    gp2 += gJ[0][0] * (-cam.fx * ipz2) + gJ[1][1] * (-cam.fy * ipz2) + gJ[0][2] * (2.0f * cam.fx * g.tx * ipz3) +
           gJ[1][2] * (2.0f * cam.fy * g.ty * ipz3);
    const float gtx = gJ[0][2] * (-cam.fx * ipz2), gty = gJ[1][2] * (-cam.fy * ipz2);
    if (!g.clampX) gp0 += gtx; else gp2 += gtx * (g.tx * ipz);
    if (!g.clampY) gp1 += gty; else gp2 += gty * (g.ty * ipz);
    for (int j = 0; j < 3; ++j) gm[j] = cam.R[j] * gp0 + cam.R[3 + j] * gp1 + cam.R[6 + j] * gp2;
    // SH colour, clamped at zero: d sh_k = basis_k d rgb; d dir through the basis.
    const float grgb[3] = {g.rgbRaw.x > 0.0f ? gs[6] : 0.0f, g.rgbRaw.y > 0.0f ? gs[7] : 0.0f,
                           g.rgbRaw.z > 0.0f ? gs[8] : 0.0f};
    // The direction gradient through the basis, coefficient by coefficient (each coefficient's
    // three values are used as soon as they are read); every coefficient is read before any SH
    // gradient is written below (gsh may alias shi).
    float gdir[3] = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        if (k < g.shCount) {
            float gbk = 0.0f;
            for (int c = 0; c < 3; ++c) gbk += grgb[c] * shi[k * 3 + c];
            float row[3];
            shBasisGradRow(k, g.dir[0], g.dir[1], g.dir[2], row);
            for (int j = 0; j < 3; ++j) gdir[j] += gbk * row[j];
        }
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        if (k < g.shCount) {
            for (int c = 0; c < 3; ++c) {
                if constexpr (ACC) gsh[k * 3 + c] += grgb[c] * g.basis[k];
                else gsh[k * 3 + c] = grgb[c] * g.basis[k];
            }
        }
    }
    if constexpr (!ACC)
        for (int j = g.shCount * 3; j < cam.shCoeffs * 3; ++j) gsh[j] = 0.0f;
    // dir = (m - eye) / |m - eye|.
    const float dd = gdir[0] * g.dir[0] + gdir[1] * g.dir[1] + gdir[2] * g.dir[2];
    const float idlen = 1.0f / g.dlen;
    for (int j = 0; j < 3; ++j) gm[j] += (gdir[j] - dd * g.dir[j]) * idlen;
}

// The view-independent part: from dL/dSigma and dL/dopacity (go) to the gradients of the
// log-scales (or, with g.act, the standard deviations), the raw quaternion and the opacity logit
// (or the opacity). Linear in (gSigma, go), so a sum over views may be passed in at once.
__host__ __device__ inline void gaussAdjointShape(const GaussProj& g, float splatScale, const float gSigma[3][3],
                                                  float go, float gls[3], float gq[4], float* glogit) {
    // Sigma = M M^T, M = Rq diag(s), s = exp(log-scale) * splatScale.
    float gRq[3][3], gsc[3] = {0, 0, 0};
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            const float gM = 2.0f * (gSigma[r][0] * g.M[0][c] + gSigma[r][1] * g.M[1][c] + gSigma[r][2] * g.M[2][c]);
            gRq[r][c] = gM * g.s[c];
            gsc[c] += gM * g.Rq[r][c];
        }
    // s = exp(log-scale) * splatScale, or (act) scale * splatScale.
    for (int j = 0; j < 3; ++j) gls[j] = gsc[j] * (g.act ? splatScale : g.s[j]);
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
    // opacity = sigmoid(logit), or (act) the opacity itself.
    *glogit = g.act ? go : go * g.opacity * (1.0f - g.opacity);
}

// Adjoint of projectGaussian (raw parameters): from gs = dL/d{u, v, A, B, C, opacity, r, g, b}
// to the gradients of the mean (gm), log-scales (gls), raw quaternion (gq), opacity logit and the
// SH coefficients (gsh, [coefficient][channel], zero past the evaluated degree). shi holds the
// Gaussian's SH, [coefficient][channel]; gsh may be the same array unless ACC. With ACC the SH
// gradients are added to gsh (which then holds a running sum over views) instead of stored.
// With g.act the scales and the opacity are the activated ones: gls is then the gradient w.r.t.
// the standard deviations and *glogit the one w.r.t. the opacity.
template <bool ACC = false>
__host__ __device__ inline void gaussAdjoint(const GaussProj& g, const CamParams& cam, const float gs[9],
                                             const float* shi, float gm[3], float gls[3], float gq[4], float* glogit,
                                             float* gsh) {
    float gSigma[3][3];
    gaussAdjointView<ACC>(g, cam, gs, shi, gm, gSigma, gsh);
    gaussAdjointShape(g, cam.splatScale, gSigma, gs[5], gls, gq, glogit);
}

// Pair slots in Gaussian index order: per block of 256 Gaussians, the sum of their tile counts.
// Batched views: blockIdx.y is the view; tile counts at view * gaussStride, block sums at
// view * blockStride (a single frame: gridDim.y = 1).
[[maybe_unused]] static __global__ __launch_bounds__(256) void k_slotBlockSums(uint32_t n, const uint32_t* __restrict__ tilesTouched,
                                                              uint32_t* __restrict__ blockSum, uint32_t gaussStride,
                                                              uint32_t blockStride) {
    __shared__ uint32_t sTmp[256 / kWave];
    tilesTouched += static_cast<size_t>(blockIdx.y) * gaussStride;
    blockSum += static_cast<size_t>(blockIdx.y) * blockStride;
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    uint32_t total;
    blockExclusiveScan<256>(i < n ? tilesTouched[i] : 0u, sTmp, &total);
    if (threadIdx.x == 0) blockSum[blockIdx.x] = total;
}

// One workgroup per view (kScanBlock wide, one round of loads up to 4096 blocks): exclusive scan
// of the per-block sums in place.
[[maybe_unused]] static __global__ __launch_bounds__(kScanBlock) void k_scanBlockSums(uint32_t* __restrict__ blockSum, uint32_t blocks,
                                                              uint32_t blockStride) {
    __shared__ uint32_t sTmp[kScanBlock / kWave];
    blockSum += static_cast<size_t>(blockIdx.y) * blockStride;
    constexpr int kPer = 4;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < blocks; base += kScanBlock * kPer) {
        const uint32_t first = base + threadIdx.x * kPer;
        uint32_t v[kPer], sum = 0;
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            v[k] = first + k < blocks ? blockSum[first + k] : 0u;
            sum += v[k];
        }
        uint32_t total;
        uint32_t run = carry + blockExclusiveScan<kScanBlock>(sum, sTmp, &total);
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            if (first + k < blocks) blockSum[first + k] = run;
            run += v[k];
        }
        carry += total;
    }
}

// slotStart[i]: the first pair slot of Gaussian i (exclusive scan of tile counts in index order).
[[maybe_unused]] static __global__ __launch_bounds__(256) void k_slotStart(uint32_t n, const uint32_t* __restrict__ tilesTouched,
                                                          const uint32_t* __restrict__ blockBase,
                                                          uint32_t* __restrict__ slotStart, uint32_t gaussStride,
                                                          uint32_t blockStride) {
    __shared__ uint32_t sTmp[256 / kWave];
    tilesTouched += static_cast<size_t>(blockIdx.y) * gaussStride;
    slotStart += static_cast<size_t>(blockIdx.y) * gaussStride;
    blockBase += static_cast<size_t>(blockIdx.y) * blockStride;
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    uint32_t total;
    const uint32_t excl = blockExclusiveScan<256>(i < n ? tilesTouched[i] : 0u, sTmp, &total);
    if (i < n) slotStart[i] = blockBase[blockIdx.x] + excl;
}

constexpr int kShRow = 49;  // LDS row stride of one Gaussian's SH (odd: row-per-lane reads are conflict-free)

// Row of flat element e in rows of k3 floats, e < 64 * 48: (e + 0.5) / k3 lies at least 0.5 / 48
// from an integer and the float product is within 2e-4 of it, so truncation gives e / k3.
__device__ inline int rowOf(int e, float rcpK3) {
    return static_cast<int>((static_cast<float>(e) + 0.5f) * rcpK3);
}

// One step of a segmented inclusive scan of x over the lanes (the steps of waveInclusive): a lane
// adds the DPP source lane's partial sums when that lane has the same owner (owners are
// non-decreasing over the lanes, so the lanes in between have it too).
template <int CTRL, int ROWS = 0xf>
__device__ inline void segmentStep(int owner, float (&x)[9]) {
    const bool same = dppMove<CTRL, ROWS>(-2, owner) == owner;  // owners are at least -1
#pragma unroll
    for (int c = 0; c < 9; ++c) {
        const float y = __int_as_float(dppMove<CTRL, ROWS>(0, __float_as_int(x[c])));
        x[c] = same ? x[c] + y : x[c];
    }
}

// Record sums of one wavefront's 64 Gaussians (lane = Gaussian, slots [s0, s0 + tt)) over their
// records [r0, r1), 64 records per chunk, one record per lane: the owner of each record comes
// from head flags at the owners' first slots and a wavefront max-scan (as in the pair
// expansion); a segmented inclusive scan over the lanes sums each owner's records within the
// chunk, and the lane ending a segment adds it to the owner's accumulator in LDS. Every lane does
// the same work whatever its Gaussian's record count. The chunk's records arrive in a0..a2 (this
// lane's record c0 + lane, zero past r1), and the next chunk's are loaded while this one is summed.
// Returns this lane's sums in gs. The order of additions is fixed, so results are deterministic.
struct RecordScratch {
    float acc[kWave * 9];
    int head[kWave];
};

__device__ inline void sumRecords(const float4* __restrict__ rec, uint32_t r0, uint32_t r1, uint32_t s0,
                                  uint32_t tt, float4 a0, float4 a1, float4 a2, RecordScratch& S, float gs[9]) {
    const int lane = __lane_id();
#pragma unroll
    for (int c = 0; c < 9; ++c) S.acc[lane * 9 + c] = 0.0f;
    int carry = -1;
    for (uint32_t c0 = r0; c0 < r1; c0 += kWave) {
        float x[9] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w, a2.x};
        const uint32_t nx = c0 + kWave + lane;
        if (nx < r1) {
            const size_t r = static_cast<size_t>(nx) * 3;
            a0 = rec[r];
            a1 = rec[r + 1];
            a2 = rec[r + 2];
        }
        waveSync();
        S.head[lane] = -1;
        waveSync();
        if (tt > 0 && s0 >= c0 && s0 < c0 + kWave) S.head[s0 - c0] = lane;
        waveSync();
        const int owner = max(carry, waveInclusiveMax(S.head[lane]));
        carry = __builtin_amdgcn_readlane(owner, kWave - 1);
        const bool has = c0 + lane < r1;
        segmentStep<kDppRowShr1>(owner, x);
        segmentStep<kDppRowShr2>(owner, x);
        segmentStep<kDppRowShr4>(owner, x);
        segmentStep<kDppRowShr8>(owner, x);
        segmentStep<kDppRowBcast15, 0xa>(owner, x);
        segmentStep<kDppRowBcast31, 0xc>(owner, x);
        const int next = dppMove<kDppWaveShl1>(-2, owner);
        const bool end = has && (lane == kWave - 1 || c0 + lane + 1 >= r1 || next != owner);
        if (end)
#pragma unroll
            for (int c = 0; c < 9; ++c) S.acc[owner * 9 + c] += x[c];
    }
    waveSync();
#pragma unroll
    for (int c = 0; c < 9; ++c) gs[c] = S.acc[lane * 9 + c];
}

// Per Gaussian over several views (the batched differentiable frames), one wavefront per 64
// consecutive Gaussians (index order). For each view in turn (v = 0, 1, ...) the wavefront's pair records of
// that view are one contiguous run [slotStart of its first Gaussian, end of its last) starting at
// pair pairBase[v] of `partial` (pairBase null: 0); sumRecords adds them up per Gaussian, 64
// records at a time (the next chunk prefetched while the current one is summed). The lane then projects its Gaussian
// into the view again and applies the view-dependent part of the adjoint (gaussAdjointView): the
// mean and SH gradients, dL/dSigma and dL/dopacity are added in view order into running sums in
// registers. The wavefront's 64 SH rows are loaded once, coalesced, into LDS (12.5 KB), where every
// view's projection and adjoint read them (read from global memory per view, they missed the L2
// and were fetched again for each view). View v + 1's first loads are issued while view v is
// processed.
// The view-independent part (gaussAdjointShape, linear in dL/dSigma and dL/dopacity) runs once on
// the sums. The SH gradients go out through LDS so that global stores are coalesced. Gaussians
// culled in every view get zero gradients. P.screenGrads (if set): dL/d(u, v) per view [view][n].
__device__ inline void gaussBackwardWave(int n, int views, const DiffParamsDev& P, const CamParams* __restrict__ cams,
                                         const uint32_t* __restrict__ tilesTouched,
                                         const uint32_t* __restrict__ slotStart,
                                         const uint64_t* __restrict__ pairBase, const float* __restrict__ partial) {
    __shared__ RecordScratch sRecSum;
    __shared__ float sSh[kWave * kShRow];
    const int lane = threadIdx.x;
    const int g0 = blockIdx.x * kWave;
    const int rows = min(kWave, n - g0);
    const int i = g0 + lane;
    const bool valid = lane < rows;
    const int K3 = P.shCoeffs * 3;  // at most 48
    const int total = rows * K3;
    const float rcpK3 = 1.0f / static_cast<float>(K3);
    {
        // The 64 rows are one contiguous block: every load issued first, then one row per lane.
        const float* src = P.sh + static_cast<size_t>(g0) * K3;
        float v[48];
#pragma unroll
        for (int j = 0; j < 48; ++j) {
            const int e = lane + j * kWave;
            v[j] = e < total ? src[e] : 0.0f;
        }
#pragma unroll
        for (int j = 0; j < 48; ++j) {
            const int e = lane + j * kWave;
            if (e < total) {
                const int r = rowOf(e, rcpK3);
                sSh[r * kShRow + (e - r * K3)] = v[j];
            }
        }
    }
    float* shRow = sSh + lane * kShRow;
    float gsh[48];
#pragma unroll
    for (int j = 0; j < 48; ++j) gsh[j] = 0.0f;
    waveSync();
    F3 m{0.0f, 0.0f, 0.0f};
    float logit = 0.0f;
    GaussProj g;
    if (valid) {
        m = reinterpret_cast<const F3*>(P.pos)[i];
        logit = P.opacity[i];
        prepareGaussian<true>(P.rot[i], reinterpret_cast<const F3*>(P.scale)[i], cams[0].splatScale, g, P.act != 0);
    }
    float gm[3] = {0, 0, 0}, gSig[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, go = 0.0f;
    bool any = false;
    const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    // Software pipeline over the views: view v + 1's tile count, first slot and first chunk of
    // records are loaded while view v is processed, so that no view starts with two dependent
    // round trips to memory.
    auto recOf = [&](int v) {
        return reinterpret_cast<const float4*>(partial + (pairBase ? pairBase[v] : uint64_t(0)) * kPartialStride);
    };
    uint32_t ttN = valid ? tilesTouched[i] : 0u;
    uint32_t s0N = valid ? slotStart[i] : 0u;
    float4 n0 = zero, n1 = zero, n2 = zero;
    {
        const uint32_t r0 = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(s0N)));
        const uint32_t r1 = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(s0N + ttN), rows - 1));
        if (r0 + lane < r1) {
            const float4* rec = recOf(0);
            const size_t r = static_cast<size_t>(r0 + lane) * 3;
            n0 = rec[r];
            n1 = rec[r + 1];
            n2 = rec[r + 2];
        }
    }
    for (int v = 0; v < views; ++v) {
        const size_t vo = static_cast<size_t>(v) * n;
        const uint32_t tt = ttN, s0 = s0N;
        float4 a0 = n0, a1 = n1, a2 = n2;
        if (v + 1 < views) {
            ttN = valid ? tilesTouched[vo + n + i] : 0u;
            s0N = valid ? slotStart[vo + n + i] : 0u;
        }
        const uint32_t r0 = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(s0)));
        const uint32_t r1 = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(s0 + tt), rows - 1));
        const float4* rec = recOf(v);
        float gs[9];
        sumRecords(rec, r0, r1, s0, tt, a0, a1, a2, sRecSum, gs);
        if (v + 1 < views) {
            // The next view's first chunk, in flight during this view's adjoint.
            n0 = n1 = n2 = zero;
            const uint32_t q0 = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(s0N)));
            const uint32_t q1 = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(s0N + ttN), rows - 1));
            if (q0 + lane < q1) {
                const float4* recN = recOf(v + 1);
                const size_t r = static_cast<size_t>(q0 + lane) * 3;
                n0 = recN[r];
                n1 = recN[r + 1];
                n2 = recN[r + 2];
            }
        }
        if (valid && P.screenGrads) P.screenGrads[vo + i] = make_float2(gs[0], gs[1]);
        if (tt > 0 && cameraTransform(m, cams[v], g) &&
            finishProjection<false, true, false, true>(m, logit, shRow, 1, cams[v], g)) {
            float vm[3], vSig[3][3];
            gaussAdjointView<true>(g, cams[v], gs, shRow, vm, vSig, gsh);
            for (int j = 0; j < 3; ++j) gm[j] += vm[j];
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) gSig[r][c] += vSig[r][c];
            go += gs[5];
            any = true;
        }
    }
    if (valid) {
        float gls[3] = {0, 0, 0}, gq[4] = {0, 0, 0, 0}, gl = 0.0f;
        if (any) {
            // The rotation and scale matrices are formed again here so that they need not stay in
            // registers through the view loop.
            GaussProj h;
            prepareGaussian<true>(P.rot[i], reinterpret_cast<const F3*>(P.scale)[i], cams[0].splatScale, h,
                                  P.act != 0);
            h.opacity = h.act ? logit : 1.0f / (1.0f + expf(-logit));  // as finishProjection sets it
            gaussAdjointShape(h, cams[0].splatScale, gSig, go, gls, gq, &gl);
        }
        for (int j = 0; j < 3; ++j) P.gMeans[3 * i + j] = gm[j], P.gScales[3 * i + j] = gls[j];
        reinterpret_cast<float4*>(P.gQuats)[i] = make_float4(gq[0], gq[1], gq[2], gq[3]);
        P.gOpacity[i] = gl;
    }
    // The SH gradients go out through LDS (the SH rows are no longer needed) so that the global
    // stores are coalesced.
    waveSync();
#pragma unroll
    for (int j = 0; j < 48; ++j)
        if (j < K3) shRow[j] = gsh[j];
    waveSync();
    float* dst = P.gSh + static_cast<size_t>(g0) * K3;
#pragma unroll
    for (int j = 0; j < 48; ++j) {
        const int e = lane + j * kWave;
        if (e < total) {
            const int r = rowOf(e, rcpK3);
            dst[e] = sSh[r * kShRow + (e - r * K3)];
        }
    }
}

// Per Gaussian, one wavefront per 64 consecutive Gaussians (index order). The wavefront's pair
// records are one contiguous run [slotStart of its first Gaussian, end of its last); sumRecords
// adds them up per Gaussian, 64 records at a time (the next chunk prefetched while the current
// one is summed).
// The SH of the 64 Gaussians is one contiguous block: it is loaded coalesced (every load issued
// up front), transposed through LDS to one row per lane, and the SH gradients go back the same
// way. gaussAdjoint then gives the gradients of means, log-scales, quaternions (w, x, y, z),
// opacity logits and SH [coefficient][channel] (with act: standard deviations and opacities in
// place of log-scales and logits); the summed splat-record gradient goes to splatGrad. Culled
// Gaussians get zero gradients. (Batched frames use gaussBackwardWave, which adds the views.)
[[maybe_unused]] static __global__ __launch_bounds__(kWave) void k_gaussBackward(
    int n, const float* __restrict__ pos, const float* __restrict__ logScale, const float4* __restrict__ quat,
    const float* __restrict__ logit, const float* __restrict__ sh, CamParams cam,
    const uint32_t* __restrict__ tilesTouched, const uint32_t* __restrict__ slotStart,
    const float* __restrict__ partial, float* __restrict__ splatGrad, float* __restrict__ gMeans,
    float* __restrict__ gScales, float* __restrict__ gQuats, float* __restrict__ gLogit, float* __restrict__ gSh,
    const FrameCtl* __restrict__ ctl, bool act) {
    __shared__ RecordScratch sRecSum;
    __shared__ float sSh[kWave * kShRow];
    if (ctl->overflow) return;
    const int lane = threadIdx.x;
    const int g0 = blockIdx.x * kWave;
    const int rows = min(kWave, n - g0);
    const int i = g0 + lane;
    const bool valid = lane < rows;
    const uint32_t tt = valid ? tilesTouched[i] : 0u;
    const uint32_t s0 = valid ? slotStart[i] : 0u;
    const int K3 = cam.shCoeffs * 3;  // at most 48
    const int total = rows * K3;
    const float* src = sh + static_cast<size_t>(g0) * K3;
    float v[48];
#pragma unroll
    for (int j = 0; j < 48; ++j) {
        const int e = lane + j * kWave;
        v[j] = e < total ? src[e] : 0.0f;
    }
    const uint32_t r0 = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(s0)));
    const uint32_t r1 = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(s0 + tt), rows - 1));
    const float4* rec = reinterpret_cast<const float4*>(partial);
    const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float4 a0 = zero, a1 = zero, a2 = zero;
    if (r0 + lane < r1) {
        const size_t r = static_cast<size_t>(r0 + lane) * 3;
        a0 = rec[r];
        a1 = rec[r + 1];
        a2 = rec[r + 2];
    }
    const float rcpK3 = 1.0f / static_cast<float>(K3);
#pragma unroll
    for (int j = 0; j < 48; ++j) {
        const int e = lane + j * kWave;
        if (e < total) {
            const int row = rowOf(e, rcpK3);
            sSh[row * kShRow + (e - row * K3)] = v[j];
        }
    }

    float gs[9];
    sumRecords(rec, r0, r1, s0, tt, a0, a1, a2, sRecSum, gs);
    waveSync();
    float* row = sSh + lane * kShRow;
    if (valid) {
        for (int c = 0; c < 9; ++c) splatGrad[static_cast<size_t>(i) * 9 + c] = gs[c];
        GaussProj g;
        if (tt > 0 && projectGaussian<false, true, false, true>(i, pos, logScale, quat, logit, row, 1, cam, g, act)) {
            float gm[3], gls[3], gq[4], gl, gSigma[3][3];
            const float opacity = act ? logit[i] : 1.0f / (1.0f + expf(-logit[i]));  // as the forward's
            gaussAdjointView<false>(g, cam, gs, row, gm, gSigma, row);
            // The rotation and scale matrices are formed again for the shape part, so that they
            // need not stay in registers through the view-dependent part.
            GaussProj h;
            prepareGaussian<true>(quat[i], reinterpret_cast<const F3*>(logScale)[i], cam.splatScale, h, act);
            h.opacity = opacity;
            gaussAdjointShape(h, cam.splatScale, gSigma, gs[5], gls, gq, &gl);
            for (int j = 0; j < 3; ++j) gMeans[3 * i + j] = gm[j], gScales[3 * i + j] = gls[j];
            reinterpret_cast<float4*>(gQuats)[i] = make_float4(gq[0], gq[1], gq[2], gq[3]);
            gLogit[i] = gl;
        } else {
            for (int j = 0; j < 3; ++j) gMeans[3 * i + j] = 0.0f, gScales[3 * i + j] = 0.0f;
            reinterpret_cast<float4*>(gQuats)[i] = zero;
            gLogit[i] = 0.0f;
            for (int j = 0; j < K3; ++j) row[j] = 0.0f;
        }
    }
    waveSync();
    float* dst = gSh + static_cast<size_t>(g0) * K3;
#pragma unroll
    for (int j = 0; j < 48; ++j) {
        const int e = lane + j * kWave;
        if (e < total) {
            const int r = rowOf(e, rcpK3);
            dst[e] = sSh[r * kShRow + (e - r * K3)];
        }
    }
}

}  // namespace ans
