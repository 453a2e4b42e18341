// Forward rasterizer kernels for gfx950: frame init, per-Gaussian preprocess, pair expansion,
// tile ranges and the tile blend. The library is compiled with -ffp-contract=off, so every
// a * b + c in the preprocess rounds twice, exactly like the CPU oracle; the blend spells out
// its fused multiply-adds explicitly.
#pragma once

#include "common.hpp"
#include "sort.hpp"

namespace ans {

struct F3 {
    float x, y, z;
};

// Real SH constants (src/sh.cpp:10-14; closed forms rounded to float).
constexpr float kShC1 = 0.4886025119029199f;
constexpr float kShC2[5] = {1.0925484305920792f, -1.0925484305920792f, 0.31539156525252005f, -1.0925484305920792f,
                               0.5462742152960396f};
constexpr float kShC3[7] = {-0.5900435899266435f, 2.890611442640554f, -0.4570457994644658f, 0.3731763325901154f,
                               -0.4570457994644658f, 1.445305721320277f, -0.5900435899266435f};

// Basis values of the unit direction (x, y, z) up to `degree`, in the order the coefficients are
// stored (src/sh.cpp:23-51, same operation order); returns the coefficient count.
__host__ __device__ inline int shBasis(int degree, float x, float y, float z, float b[16]) {
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
    return count;
}

// 0.5 + sum_k b_k sh_k per channel (before the clamp at zero; src/sh.cpp:53-55). Coefficient
// j = k * 3 + channel sits at sh[j * stride]: stride 1 for one Gaussian's [coefficient][channel]
// block, stride n for the planar device layout [coefficient * 3 + channel][n]. EAGER (SH in
// global memory): all 48 coefficients of degree 3 are loaded before the first is used; without it
// (SH in LDS, where a read's latency is short) each coefficient's three values are read as they
// are used, so that the 48 do not occupy registers at once.
template <bool PLANAR, bool EAGER = true>
__host__ __device__ inline float3 shAccumulate(const float* __restrict__ sh, size_t stride, int count, const float b[16]) {
    float3 c = make_float3(0.5f, 0.5f, 0.5f);
    if (EAGER && count == 16) {
        float v[48];
        if constexpr (PLANAR) {
#pragma unroll
            for (int j = 0; j < 48; ++j) v[j] = sh[j * stride];
        } else {
            const float4* s4 = reinterpret_cast<const float4*>(sh);
#pragma unroll
            for (int q = 0; q < 12; ++q) {
                const float4 t = s4[q];
                v[4 * q] = t.x; v[4 * q + 1] = t.y; v[4 * q + 2] = t.z; v[4 * q + 3] = t.w;
            }
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
                c.x += b[k] * sh[(k * 3 + 0) * stride];
                c.y += b[k] * sh[(k * 3 + 1) * stride];
                c.z += b[k] * sh[(k * 3 + 2) * stride];
            }
        }
    }
    return c;
}

__host__ __device__ inline float clampf(float v, float lo, float hi) { return (v < lo) ? lo : (hi < v) ? hi : v; }

// One projected Gaussian and the intermediates its adjoint reuses.
struct GaussProj {
    float u, v, A, B, C, opacity, depth;
    float3 rgb, rgbRaw;
    int x0, y0, x1, y1;
    float px, py, pz, tx, ty;
    bool clampX, clampY;
    float s[3], q[4], qlen;
    bool act;  // RAW with activated scales and opacities (standard deviations, opacities in [0, 1])
    float Rq[3][3], M[3][3], Sigma[3][3], T[2][3];
    float a, b, c, det;
    float dir[3], dlen;
    float basis[16];
    int shCount;
    int radius;  // ceil(3 sqrt(lambda_max)), pixels
};

// Projection of one Gaussian (the maths of src/render.cpp:39-104 with the FP32 camera transform
// of src/gpu_render.cu:94-96), in three parts below: world -> camera; near cull; Sigma = M M^T, M = R(q) diag(s); EWA 2D
// covariance with the clamped Jacobian and the 0.3 low-pass; conic; radius = ceil(3 sqrt(lambda));
// centre in pixels; screen cull; tile rect; SH colour. RAW: the parameters are the optimizer's
// (scale = exp(log-scale), opacity = sigmoid(logit), rotation = q / |q|, src/ply_loader.cpp:124-130),
// activated here in FP32. Deliberate deviations shared with the CPU oracle: the near test is
// !(z >= 0.2) so a NaN depth is culled, and a non-finite centre or a radius of 2^24 pixels or more
// is culled.

// World -> camera (FP32, left to right as gpu_render.cu:94-96) and the near cull; false if culled.
__host__ __device__ inline bool cameraTransform(const F3 m, const CamParams& cam, GaussProj& g) {
    g.px = cam.R[0] * m.x + cam.R[1] * m.y + cam.R[2] * m.z + cam.t[0];
    g.py = cam.R[3] * m.x + cam.R[4] * m.y + cam.R[5] * m.z + cam.t[1];
    g.pz = cam.R[6] * m.x + cam.R[7] * m.y + cam.R[8] * m.z + cam.t[2];
    return g.pz >= 0.2f;
}

// The view-independent part: rotation, scale (times splatScale), M = R(q) diag(s), Sigma = M M^T.
// RAW with act: the quaternion is normalized but the scales and the opacity are taken as given
// (the convention of rasterizers that receive exp(log-scale) and sigmoid(logit) from the caller,
// pipeline/train_gsplat.py:98-101).
template <bool RAW>
__host__ __device__ inline void prepareGaussian(const float4 q4, const F3 sc, float splatScale, GaussProj& g,
                                                bool act = false) {
    g.act = RAW && act;
    if constexpr (RAW) {
        g.qlen = sqrtf(q4.x * q4.x + q4.y * q4.y + q4.z * q4.z + q4.w * q4.w);
        if (g.qlen > 0.0f) {
            g.q[0] = q4.x / g.qlen; g.q[1] = q4.y / g.qlen; g.q[2] = q4.z / g.qlen; g.q[3] = q4.w / g.qlen;
        } else {
            g.q[0] = 1.0f; g.q[1] = g.q[2] = g.q[3] = 0.0f;
        }
    } else {
        g.qlen = 1.0f;
        g.q[0] = q4.x; g.q[1] = q4.y; g.q[2] = q4.z; g.q[3] = q4.w;
    }
    const float w = g.q[0], x = g.q[1], y = g.q[2], z = g.q[3];
    g.Rq[0][0] = 1 - 2 * (y * y + z * z); g.Rq[0][1] = 2 * (x * y - w * z);     g.Rq[0][2] = 2 * (x * z + w * y);
    g.Rq[1][0] = 2 * (x * y + w * z);     g.Rq[1][1] = 1 - 2 * (x * x + z * z); g.Rq[1][2] = 2 * (y * z - w * x);
    g.Rq[2][0] = 2 * (x * z - w * y);     g.Rq[2][1] = 2 * (y * z + w * x);     g.Rq[2][2] = 1 - 2 * (x * x + y * y);
    if (RAW && !g.act) {
        g.s[0] = expf(sc.x) * splatScale; g.s[1] = expf(sc.y) * splatScale; g.s[2] = expf(sc.z) * splatScale;
    } else {
        g.s[0] = sc.x * splatScale; g.s[1] = sc.y * splatScale; g.s[2] = sc.z * splatScale;
    }
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) g.M[r][c] = g.Rq[r][c] * g.s[c];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) g.Sigma[r][c] = g.M[r][0] * g.M[c][0] + g.M[r][1] * g.M[c][1] + g.M[r][2] * g.M[c][2];
}

// Unit direction from the camera centre to the Gaussian, and the distance.
__host__ __device__ inline void setViewDirection(const F3 m, const CamParams& cam, GaussProj& g) {
    g.dir[0] = m.x - cam.eye[0];
    g.dir[1] = m.y - cam.eye[1];
    g.dir[2] = m.z - cam.eye[2];
    g.dlen = sqrtf(g.dir[0] * g.dir[0] + g.dir[1] * g.dir[1] + g.dir[2] * g.dir[2]);
    g.dir[0] = g.dir[0] / g.dlen;
    g.dir[1] = g.dir[1] / g.dlen;
    g.dir[2] = g.dir[2] / g.dlen;
}

template <bool PLANAR, bool RAW, bool EAGER_SH = true, bool ADJ = false>
__host__ __device__ inline bool finishProjection(const F3 m, float opacityIn, const float* __restrict__ shBase,
                                                 size_t shStride, const CamParams& cam, GaussProj& g);

// Projects Gaussian i: cameraTransform, prepareGaussian, finishProjection. shBase points at this
// Gaussian's coefficient 0 (red); coefficient * 3 + channel j sits at shBase[j * shStride]
// (PLANAR: shStride = n; otherwise 1 and shBase 16-byte aligned).
template <bool PLANAR, bool RAW, bool EAGER_SH = true, bool ADJ = false>
__host__ __device__ inline bool projectGaussian(int i, const float* __restrict__ pos, const float* __restrict__ scale,
                                                const float4* __restrict__ rot, const float* __restrict__ opacity,
                                                const float* __restrict__ shBase, size_t shStride, const CamParams& cam,
                                                GaussProj& g, bool act = false) {
    const F3 m = reinterpret_cast<const F3*>(pos)[i];
    if (!cameraTransform(m, cam, g)) return false;
    prepareGaussian<RAW>(rot[i], reinterpret_cast<const F3*>(scale)[i], cam.splatScale, g, act);
    return finishProjection<PLANAR, RAW, EAGER_SH, ADJ>(m, opacity[i], shBase, shStride, cam, g);
}

// The view-dependent rest: EWA 2D covariance with the clamped Jacobian and the 0.3 low-pass,
// conic, radius, centre, screen cull, tile rect, SH colour, opacity (sigmoid of opacityIn when
// RAW). Needs cameraTransform and prepareGaussian done on g. ADJ (the adjoint of a Gaussian the
// forward kept): only what the adjoint reads (the covariance and its Jacobian, the view
// direction, SH basis and unclamped colour), with the same operations; the radius, centre,
// screen cull, tile rect, conic and opacity are skipped.
template <bool PLANAR, bool RAW, bool EAGER_SH, bool ADJ>
__host__ __device__ inline bool finishProjection(const F3 m, float opacityIn, const float* __restrict__ shBase,
                                                 size_t shStride, const CamParams& cam, GaussProj& g) {
    const float px = g.px, py = g.py, pz = g.pz;
    const float rx = px / pz, ry = py / pz;
    g.clampX = rx < -cam.limX || cam.limX < rx;
    g.clampY = ry < -cam.limY || cam.limY < ry;
    g.tx = clampf(rx, -cam.limX, cam.limX) * pz;
    g.ty = clampf(ry, -cam.limY, cam.limY) * pz;
    const float J[2][3] = {{cam.fx / pz, 0, -cam.fx * g.tx / (pz * pz)}, {0, cam.fy / pz, -cam.fy * g.ty / (pz * pz)}};
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) g.T[r][c] = J[r][0] * cam.R[c] + J[r][1] * cam.R[3 + c] + J[r][2] * cam.R[6 + c];
    float TS[2][3];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) TS[r][c] = g.T[r][0] * g.Sigma[0][c] + g.T[r][1] * g.Sigma[1][c] + g.T[r][2] * g.Sigma[2][c];
    g.a = TS[0][0] * g.T[0][0] + TS[0][1] * g.T[0][1] + TS[0][2] * g.T[0][2] + 0.3f;
    g.b = TS[0][0] * g.T[1][0] + TS[0][1] * g.T[1][1] + TS[0][2] * g.T[1][2];
    g.c = TS[1][0] * g.T[1][0] + TS[1][1] * g.T[1][1] + TS[1][2] * g.T[1][2] + 0.3f;
    g.det = g.a * g.c - g.b * g.b;
    if (!(g.det > 0.0f)) return false;

    if constexpr (ADJ) {
        setViewDirection(m, cam, g);
        g.shCount = shBasis(cam.shDegree, g.dir[0], g.dir[1], g.dir[2], g.basis);
        g.rgbRaw = shAccumulate<PLANAR, EAGER_SH>(shBase, shStride, g.shCount, g.basis);
        return true;
    }
    const float mid = 0.5f * (g.a + g.c);
    const float disc = mid * mid - g.det;
    const float lambda = mid + sqrtf((0.1f < disc) ? disc : 0.1f);
    const float rf = ceilf(3.0f * sqrtf(lambda));
    if (!(rf < 16777216.0f)) return false;
    const int radius = static_cast<int>(rf);
    g.radius = radius;
    g.u = cam.fx * px / pz + cam.cx;
    g.v = cam.fy * py / pz + cam.cy;
    if (!__builtin_isfinite(g.u) || !__builtin_isfinite(g.v)) return false;
    if (g.u + radius < 0 || g.v + radius < 0 || g.u - radius >= cam.width || g.v - radius >= cam.height) return false;

    // The reference divides the truncated bounds by the tile size (truncation toward zero); the
    // tile sizes are powers of two, so an arithmetic shift gives the same tile for every bound
    // that is not negative, and a negative lower bound clamps to 0 either way. The upper bounds
    // are not negative (u + radius >= 0 after the screen cull).
    g.x0 = max(0, static_cast<int>(g.u - radius) >> cam.tileShiftX);
    g.y0 = max(0, static_cast<int>(g.v - radius) >> cam.tileShiftY);
    g.x1 = min(cam.tilesX - 1, static_cast<int>(g.u + radius) >> cam.tileShiftX);
    g.y1 = min(cam.tilesY - 1, static_cast<int>(g.v + radius) >> cam.tileShiftY);

    setViewDirection(m, cam, g);
    g.shCount = shBasis(cam.shDegree, g.dir[0], g.dir[1], g.dir[2], g.basis);
    g.rgbRaw = shAccumulate<PLANAR, EAGER_SH>(shBase, shStride, g.shCount, g.basis);
    // std::max(0, v) semantics (NaN -> 0).
    g.rgb.x = (0.0f < g.rgbRaw.x) ? g.rgbRaw.x : 0.0f;
    g.rgb.y = (0.0f < g.rgbRaw.y) ? g.rgbRaw.y : 0.0f;
    g.rgb.z = (0.0f < g.rgbRaw.z) ? g.rgbRaw.z : 0.0f;
    if (RAW && !g.act) g.opacity = 1.0f / (1.0f + expf(-opacityIn));
    else g.opacity = opacityIn;
    g.A = g.c / g.det;
    g.B = -g.b / g.det;
    g.C = g.a / g.det;
    g.depth = pz;
    return true;
}

// Tile rects packed as (x0 | y0 << 16, x1 | y1 << 16), inclusive; a culled Gaussian's is the
// empty rect (1, 0) (x1 < x0). A visible Gaussian's rect is never empty: the screen cull keeps
// u - r < width and u + r >= 0, so 0 <= x0 <= x1 <= tilesX - 1, and likewise in y.
constexpr uint2 kEmptyRect = {1u, 0u};

__device__ inline uint32_t rectTileCount(uint2 rp) {
    const uint32_t x0 = rp.x & 0xffffu, y0 = rp.x >> 16, x1 = rp.y & 0xffffu, y1 = rp.y >> 16;
    return x1 >= x0 ? (x1 - x0 + 1) * (y1 - y0 + 1) : 0u;
}

// Writes a culled Gaussian's tile count (0, also in its record), empty rect and depth key (the
// largest).
__device__ inline void storeCulled(size_t o, SplatRec* __restrict__ rec, uint32_t* __restrict__ tilesTouched,
                                   uint32_t* __restrict__ depthKey, uint2* __restrict__ rect) {
    rec[o].meta = make_uint4(0u, 0u, 0u, 0u);
    if (tilesTouched) tilesTouched[o] = 0;
    rect[o] = kEmptyRect;
    depthKey[o] = 0xffffffffu;
}

// A visible Gaussian's splat record (with its tile rect and count) and its packed tile rect.
__device__ inline SplatRec splatRecord(const GaussProj& g, uint2& rp) {
    const uint32_t tt = static_cast<uint32_t>((g.x1 - g.x0 + 1) * (g.y1 - g.y0 + 1));
    rp = make_uint2(static_cast<uint32_t>(g.x0) | (static_cast<uint32_t>(g.y0) << 16),
                    static_cast<uint32_t>(g.x1) | (static_cast<uint32_t>(g.y1) << 16));
    SplatRec r;
    r.geo = make_float4(g.u, g.v, g.A, g.B);
    r.geo2 = make_float4(g.C, g.opacity, 0.0f, 0.0f);
    r.col = make_float4(g.rgb.x, g.rgb.y, g.rgb.z, g.depth);
    r.meta = make_uint4(rp.x, rp.y, tt, static_cast<uint32_t>(g.radius));
    return r;
}

// Writes a visible Gaussian's splat record, tile count (unless tilesTouched is null), packed tile
// rect (a compact copy of the record's, which the pair count gathers in depth order) and depth key
// at index o.
__device__ inline void storeSplat(const GaussProj& g, size_t o, SplatRec* __restrict__ rec,
                                  uint32_t* __restrict__ tilesTouched, uint32_t* __restrict__ depthKey,
                                  uint2* __restrict__ rect) {
    uint2 rp;
    const SplatRec r = splatRecord(g, rp);
    rec[o] = r;
    rect[o] = rp;
    if (tilesTouched) tilesTouched[o] = r.meta.z;
    depthKey[o] = __float_as_uint(g.depth);
}

// Writes the records staged in sRec (one per lane of a 256-lane workgroup) for Gaussians
// [first, first + 256) below n: lane l writes 16-byte words l, l + 256, l + 512, l + 768 of the
// run, so each store of a wavefront covers one contiguous kilobyte. A culled Gaussian's record
// (tile count 0, staged as zeros) gets only its meta word, as storeCulled writes it. Called by
// every lane after the staging and a barrier.
__device__ inline void writeStagedRecords(const SplatRec* sRec, SplatRec* __restrict__ rec, uint32_t first,
                                          uint32_t n) {
    const uint32_t words = first < n ? 4u * min(256u, n - first) : 0u;
    const float4* src = reinterpret_cast<const float4*>(sRec);
    float4* dst = reinterpret_cast<float4*>(rec + first);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t w = k * 256 + threadIdx.x;
        if (w < words && ((w & 3u) == 3u || sRec[w >> 2].meta.z != 0u)) dst[w] = src[w];
    }
}

// Per-Gaussian preprocess, one lane per Gaussian. Writes the splat record
// {u, v, A, B | C, opacity, 0, 0 | r, g, b, depth | rect, tiles, radius}, the packed tile rect, the
// tile count and the depth sort key (0xffffffff for culled Gaussians).
template <bool PLANAR, bool RAW>
__global__ __launch_bounds__(256) void k_preprocess(int n, const float* __restrict__ pos,
                                                    const float* __restrict__ scale, const float4* __restrict__ rot,
                                                    const float* __restrict__ opacity, const float* __restrict__ sh,
                                                    CamParams cam, SplatRec* __restrict__ rec,
                                                    uint32_t* __restrict__ tilesTouched,
                                                    uint32_t* __restrict__ depthKey, uint2* __restrict__ rect,
                                                    FrameCtl* __restrict__ ctl,
                                                    uint2* __restrict__ ranges, uint32_t tiles,
                                                    FrameStamps* __restrict__ stamps, bool act) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    // The frame starts here: its control block (all but sticky) and tile ranges are cleared by
    // the first max(n, tiles) lanes; no later kernel of the frame runs before this one ends.
    // Workgroup 0 is dispatched first, so its clock reading is the kernel's start.
    if (i == 0) {
        if (stamps) stamps->t[0] = wall_clock64();
        ctl->epoch += 1;
        ctl->visible = 0;
        ctl->pairs = 0;
        ctl->overflow = 0;
        ctl->finished = 0;
    }
    if (static_cast<uint32_t>(i) < tiles) ranges[i] = make_uint2(0, 0);
    // The workgroup's 256 records are staged in LDS and written out as 16 contiguous kilobytes
    // (writeStagedRecords), instead of four 16-byte stores per lane 64 bytes apart.
    __shared__ SplatRec sRec[256];
    SplatRec r{};
    if (i < n) {
        GaussProj g;
        const float* shBase = PLANAR ? sh + i : sh + static_cast<size_t>(i) * cam.shCoeffs * 3;
        uint2 rp = kEmptyRect;
        uint32_t key = 0xffffffffu;
        if (projectGaussian<PLANAR, RAW>(i, pos, scale, rot, opacity, shBase, PLANAR ? static_cast<size_t>(n) : 1,
                                         cam, g, act)) {
            r = splatRecord(g, rp);
            key = __float_as_uint(g.depth);
        }
        rect[i] = rp;
        if (tilesTouched) tilesTouched[i] = r.meta.z;
        depthKey[i] = key;
    }
    sRec[threadIdx.x] = r;
    __syncthreads();
    // The grid covers max(n, tiles) lanes (the range clear): workgroups past n write no records.
    writeStagedRecords(sRec, rec, blockIdx.x * 256u, static_cast<uint32_t>(n));
}

constexpr int kDupBlock = 256;                       // 4 wavefronts x 64 depth ranks
constexpr int kDupTile = kDupBlock;                  // depth ranks per block

// Batched views (blockIdx.y = view) offset per-Gaussian arrays by view * gaussStride, per-block
// totals by view * blockStride, pair arrays by view * pairStride and use ctl[view]; a single frame
// is gridDim.y = 1.

// Gaussians with more tiles than this are expanded by their whole wavefront, one at a time.
constexpr uint32_t kBigPairs = 32;

// Pairs and visible Gaussians per block of 256 depth ranks. Also gathers each rank's tile rect
// (8 bytes of the compact rect array) into depth order with its tile count (ttByRank,
// rectByRank; rect left unwritten for culled ranks), so that the expansion reads them with
// coalesced loads. keptCtl (or null): the depth sort dropped the culled Gaussians and kept
// keptCtl[view].sortCount ranks; the ranks after them count as culled.
[[maybe_unused]] static __global__ __launch_bounds__(kDupBlock) void k_pairTotals(const uint32_t* __restrict__ sortedIdx,
                                                                 const uint2* __restrict__ rect, uint32_t n,
                                                                 uint32_t* __restrict__ ttByRank,
                                                                 uint2* __restrict__ rectByRank,
                                                                 uint32_t* __restrict__ blockPairs,
                                                                 uint32_t* __restrict__ blockVisible,
                                                                 uint32_t gaussStride, uint32_t blockStride,
                                                                 FrameStamps* __restrict__ stamps,
                                                                 const FrameCtl* __restrict__ keptCtl) {
    __shared__ uint32_t sTmp[kDupBlock / kWave];
    if (stamps && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) stamps->t[1] = wall_clock64();
    const size_t go = static_cast<size_t>(blockIdx.y) * gaussStride;
    sortedIdx += go;
    rect += go;
    ttByRank += go;
    rectByRank += go;
    blockPairs += static_cast<size_t>(blockIdx.y) * blockStride;
    blockVisible += static_cast<size_t>(blockIdx.y) * blockStride;
    const uint32_t r = blockIdx.x * kDupTile + threadIdx.x;
    const uint32_t kept = keptCtl ? min(n, keptCtl[blockIdx.y].sortCount) : n;
    uint32_t c = 0;
    if (r < kept) {
        const uint2 rp = rect[sortedIdx[r]];
        c = rectTileCount(rp);
        ttByRank[r] = c;
        if (c) rectByRank[r] = rp;
    } else if (r < n) {
        ttByRank[r] = 0;
    }
    uint32_t total;
    blockExclusiveScan<kDupBlock>(c, sTmp, &total);
    const uint32_t visible = static_cast<uint32_t>(__syncthreads_count(c != 0));
    if (threadIdx.x == 0) {
        blockPairs[blockIdx.x] = total;
        blockVisible[blockIdx.x] = visible;
    }
}

constexpr int kScanBlock = 1024;  // k_scanPairs: one pass covers 4096 expansion blocks (1M ranks)

// One workgroup: exclusive scan of the per-block pair counts; publishes the pair total (the sum
// gpu_render.cu:324-329 reads back to the host) and the visible count, and flags a pair total
// above the pair buffers' capacity; with mirror, also to pinned host memory under number seq.
// The workgroup is wide so that a frame of up to 1M Gaussians takes one round of loads.
[[maybe_unused]] static __global__ __launch_bounds__(kScanBlock) void k_scanPairs(uint32_t* __restrict__ blockPairs,
                                                                uint32_t* __restrict__ blockVisible,
                                                                uint32_t blocks, uint32_t cap,
                                                                FrameCtl* __restrict__ ctl, uint32_t blockStride,
                                                                CountMirror* __restrict__ mirror, uint32_t seq) {
    __shared__ uint32_t sTmp[kScanBlock / kWave];
    blockPairs += static_cast<size_t>(blockIdx.y) * blockStride;
    blockVisible += static_cast<size_t>(blockIdx.y) * blockStride;
    ctl += blockIdx.y;
    if (mirror) mirror += blockIdx.y;
    constexpr int kPer = 4;
    uint64_t carry = 0;
    uint32_t visible = 0;
    for (uint32_t base = 0; base < blocks; base += kScanBlock * kPer) {
        const uint32_t first = base + threadIdx.x * kPer;
        uint32_t v[kPer];
        uint32_t sum = 0, vis = 0;
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const bool in = first + k < blocks;
            v[k] = in ? blockPairs[first + k] : 0u;
            vis += in ? blockVisible[first + k] : 0u;
            sum += v[k];
        }
        uint32_t total, visTotal;
        uint32_t run = static_cast<uint32_t>(carry) + blockExclusiveScan<kScanBlock>(sum, sTmp, &total);
        blockExclusiveScan<kScanBlock>(vis, sTmp, &visTotal);
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            if (first + k < blocks) blockPairs[first + k] = run;
            run += v[k];
        }
        carry += total;
        visible += visTotal;
    }
    if (threadIdx.x == 0) {
        const uint32_t pairs = carry > 0xffffffffull ? 0xffffffffu : static_cast<uint32_t>(carry);
        ctl->visible = visible;
        ctl->pairs = pairs;
        if (carry > cap) {
            ctl->overflow = 1;
            ctl->sticky = 1;
        }
        if (mirror) {
            mirror->pairs = pairs;
            mirror->visible = visible;
            mirror->overflow = carry > cap ? 1u : 0u;
            // Frames of one stream run in order, so a plain read-modify-write is enough.
            if (carry > cap) {
                mirror->sticky += 1u;
                mirror->stickyPairs = max(mirror->stickyPairs, pairs);
            }
            __hip_atomic_store(&mirror->seq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
}

// Inclusive max-scan across a 64-lane wavefront.
__device__ inline int waveInclusiveMax(int v) {
    return waveInclusive(v, INT_MIN, [](int a, int b) { return max(a, b); });
}

// Pair expansion in depth order: each wavefront takes 64 consecutive depth ranks and writes their
// pairs (tile id, Gaussian index) contiguously, each Gaussian's tiles row-major (the order
// gpu_render.cu:155-156 writes them), 64 pairs per step with one pair per lane. A step finds
// the Gaussian owning each slot by marking the slots where Gaussians start (LDS head flags) and
// a wavefront max-scan. Gaussians with more than kBigPairs tiles would make that walk spend many
// steps on one owner (depth order puts the nearest, largest splats next to each other), so the
// walk covers only the other Gaussians' slots, and afterwards the wavefront writes each of its
// wide Gaussians with all 64 lanes (writeRectPairs) into the slots reserved for it.

// The pairs of one Gaussian with a w x h tile rect at (x0, y0), written by a whole wavefront from
// slot `first`: lane l writes tiles l, l + 64, ... row-major, carrying (row, col) instead of
// dividing.
__device__ inline void writeRectPairs(uint32_t first, uint32_t g, uint32_t rectLo, uint32_t rectHi, int tilesX,
                                      uint32_t* __restrict__ tileOut, uint32_t* __restrict__ gaussOut) {
    const uint32_t lane = __lane_id();
    const uint32_t x0 = rectLo & 0xffffu, y0 = rectLo >> 16;
    const uint32_t w = (rectHi & 0xffffu) - x0 + 1, h = (rectHi >> 16) - y0 + 1;
    uint32_t slot = first + lane;
    uint32_t row = lane / w, col = lane - (lane / w) * w;
    const uint32_t stepRows = kWave / w, stepCols = kWave - stepRows * w;
    for (uint32_t k = lane; k < w * h; k += kWave, slot += kWave) {
        tileOut[slot] = (y0 + row) * static_cast<uint32_t>(tilesX) + x0 + col;
        gaussOut[slot] = g;
        row += stepRows;
        col += stepCols;
        if (col >= w) {
            col -= w;
            row += 1;
        }
    }
}

// tilesXView (batched views, else null) holds each view's tile-grid width.
[[maybe_unused]] static __global__ __launch_bounds__(kDupBlock) void k_expandPairs(
    const uint32_t* __restrict__ sortedIdx, const uint32_t* __restrict__ ttByRank,
    const uint2* __restrict__ rectByRank, uint32_t n, int tilesX, const uint32_t* __restrict__ blockBase,
    uint32_t* __restrict__ tileOut, uint32_t* __restrict__ gaussOut, const FrameCtl* __restrict__ ctl,
    uint32_t gaussStride, uint32_t blockStride, uint32_t pairStride, const int* __restrict__ tilesXView) {
    __shared__ uint32_t sWave[kDupBlock / kWave];
    __shared__ int sHead[kDupBlock / kWave][kWave];
    {
        const uint32_t view = blockIdx.y;
        ctl += view;
        const size_t go = static_cast<size_t>(view) * gaussStride, po = static_cast<size_t>(view) * pairStride;
        sortedIdx += go;
        ttByRank += go;
        rectByRank += go;
        blockBase += static_cast<size_t>(view) * blockStride;
        tileOut += po;
        gaussOut += po;
        if (tilesXView) tilesX = tilesXView[view];
    }
    const int wave = threadIdx.x / kWave, lane = __lane_id();
    const uint32_t r = blockIdx.x * kDupTile + threadIdx.x;
    // The ranks' counts, rects and indices and the block's first slot are loaded together with
    // the overflow flag (all in bounds either way), so the flag costs no round trip of its own.
    uint32_t g = 0, cnt = 0;
    uint2 rp = make_uint2(0, 0);
    if (r < n) {
        cnt = ttByRank[r];
        rp = rectByRank[r];
        g = sortedIdx[r];
    }
    const uint32_t blockFirst = blockBase[blockIdx.x];
    if (ctl->overflow) return;
    if (!cnt) {
        rp = make_uint2(0, 0);
        g = 0;
    }
    const bool isBig = cnt > kBigPairs;
    const uint32_t cntS = isBig ? 0u : cnt;  // slots the head-flag walk writes
    const uint32_t incl = waveInclusiveScan(cnt);
    const uint32_t excl = incl - cnt;
    const uint32_t inclS = waveInclusiveScan(cntS);
    const uint32_t exclS = inclS - cntS;
    const uint32_t waveTotal = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(incl), kWave - 1));
    const uint32_t waveSmall = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(inclS), kWave - 1));
    const uint64_t bigLanes = __ballot(isBig);
    if (lane == 0) sWave[wave] = waveTotal;
    __syncthreads();
    uint32_t out = blockFirst;
    for (int w = 0; w < wave; ++w) out += sWave[w];

    const uint32_t x0 = rp.x & 0xffffu, y0 = rp.x >> 16;
    const uint32_t w = (rp.y & 0xffffu) - x0 + 1;
    int carry = -1;
    for (uint32_t chunk = 0; chunk < waveSmall; chunk += kWave) {
        waveSync();
        sHead[wave][lane] = -1;
        waveSync();
        if (cntS && exclS >= chunk && exclS < chunk + kWave) sHead[wave][exclS - chunk] = lane;
        waveSync();
        const int owner = max(carry, waveInclusiveMax(sHead[wave][lane]));
        carry = __builtin_amdgcn_readlane(owner, kWave - 1);
        const uint32_t oStartS = static_cast<uint32_t>(__shfl(static_cast<int>(exclS), owner, kWave));
        const uint32_t oStart = static_cast<uint32_t>(__shfl(static_cast<int>(excl), owner, kWave));
        const uint32_t oX0 = static_cast<uint32_t>(__shfl(static_cast<int>(x0), owner, kWave));
        const uint32_t oY0 = static_cast<uint32_t>(__shfl(static_cast<int>(y0), owner, kWave));
        const uint32_t oW = static_cast<uint32_t>(__shfl(static_cast<int>(w), owner, kWave));
        const uint32_t oG = static_cast<uint32_t>(__shfl(static_cast<int>(g), owner, kWave));
        const uint32_t slotS = chunk + lane;
        if (slotS < waveSmall) {
            // row = k / oW with a float reciprocal and one correction each way (k <= kBigPairs).
            const uint32_t k = slotS - oStartS;
            uint32_t row = static_cast<uint32_t>(static_cast<float>(k) * __builtin_amdgcn_rcpf(static_cast<float>(oW)));
            int32_t col = static_cast<int32_t>(k) - static_cast<int32_t>(row * oW);
            if (col < 0) {
                row -= 1;
                col += static_cast<int32_t>(oW);
            } else if (col >= static_cast<int32_t>(oW)) {
                row += 1;
                col -= static_cast<int32_t>(oW);
            }
            const uint32_t slot = out + oStart + k;
            tileOut[slot] = (oY0 + row) * static_cast<uint32_t>(tilesX) + oX0 + static_cast<uint32_t>(col);
            gaussOut[slot] = oG;
        }
    }
    // The wavefront's wide Gaussians, one at a time with all 64 lanes, into the slots the scan
    // reserved for them in depth order.
    for (uint64_t wide = bigLanes; wide != 0; wide &= wide - 1) {
        const int L = __builtin_ctzll(wide);
        writeRectPairs(out + static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(excl), L)),
                       static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(g), L)),
                       static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(rp.x), L)),
                       static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(rp.y), L)), tilesX, tileOut,
                       gaussOut);
    }
}

// Start and end of each tile's run in the sorted pairs (gpu_render.cu:164-177); ranges were
// cleared to [0, 0) by the preprocess. Grid-stride over runs of 256 pairs per wavefront: lane l
// loads pairs l, l + 64, l + 128, l + 192 of the run (each load coalesced across the wavefront)
// and takes its neighbours' tiles from the adjacent lanes by DPP; the run's inner edges come from
// lane 0 or 63 of the neighbouring load (readlane), its outer edges from one load each.
[[maybe_unused]] static __global__ __launch_bounds__(256) void k_tileRanges(const uint32_t* __restrict__ sortedTile,
                                                           const FrameCtl* __restrict__ ctl,
                                                           uint2* __restrict__ ranges, uint32_t pairStride,
                                                           uint32_t tileStride) {
    ctl += blockIdx.y;
    sortedTile += static_cast<size_t>(blockIdx.y) * pairStride;
    ranges += static_cast<size_t>(blockIdx.y) * tileStride;
    if (ctl->overflow) return;
    const uint32_t total = ctl->pairs;
    const uint32_t lane = __lane_id();
    constexpr uint32_t kNone = 0xffffffffu;
    constexpr uint32_t kRun = 4 * kWave;
    // Wavefront-uniform loop (the DPP moves need every lane).
    for (uint32_t base = (blockIdx.x * 4 + threadIdx.x / kWave) * kRun; base < total;
         base += gridDim.x * 4 * kRun) {
        uint32_t t[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t p = base + k * kWave + lane;
            t[k] = p < total ? sortedTile[p] : kNone;
        }
        const uint32_t before = (lane == 0 && base > 0) ? sortedTile[base - 1] : kNone;
        const uint32_t after = (lane == kWave - 1 && base + kRun < total) ? sortedTile[base + kRun] : kNone;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t pOld = k == 0 ? before : static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(t[k - 1]), kWave - 1));
            const uint32_t nOld = k == 3 ? after : static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(t[k + 1]), 0));
            const uint32_t prev = static_cast<uint32_t>(dppMove<kDppWaveShr1>(static_cast<int>(pOld), static_cast<int>(t[k])));
            const uint32_t next = static_cast<uint32_t>(dppMove<kDppWaveShl1>(static_cast<int>(nOld), static_cast<int>(t[k])));
            const uint32_t p = base + k * kWave + lane;
            if (p < total) {
                if (p == 0 || prev != t[k]) ranges[t[k]].x = p;
                if (p == total - 1 || next != t[k]) ranges[t[k]].y = p + 1;
            }
        }
    }
}

// Smallest value of q(d) = A dx^2 + 2 B dx dy + C dy^2 (A, C > 0, AC > B^2) over the box
// dx in [dx0, dx1], dy in [dy0, dy1]: 0 if the box holds the origin, else the least of the
// four edge minima (q is convex, so its minimum over the box lies on the boundary).
// rB_C = -B / C and rB_A = -B / A are the unconstrained minimizers per unit of the other axis.
__device__ inline float quadMinOverBox(float A, float B, float C, float rB_C, float rB_A, float dx0, float dx1,
                                       float dy0, float dy1) {
    if (dx0 <= 0.0f && dx1 >= 0.0f && dy0 <= 0.0f && dy1 >= 0.0f) return 0.0f;
    // dx (A dx + 2B dy) + C dy^2 with fused multiply-adds: a culling bound, compared against a
    // threshold widened far beyond any rounding, so the operation order is free here.
    const float B2 = B + B;
    auto q = [&](float dx, float dy) {
        return __builtin_fmaf(dx, __builtin_fmaf(A, dx, B2 * dy), dy * (C * dy));
    };
    const float a = q(dx0, fminf(fmaxf(rB_C * dx0, dy0), dy1));
    const float b = q(dx1, fminf(fmaxf(rB_C * dx1, dy0), dy1));
    const float c = q(fminf(fmaxf(rB_A * dy0, dx0), dx1), dy0);
    const float d = q(fminf(fmaxf(rB_A * dy1, dx0), dx1), dy1);
    return fminf(fminf(a, b), fminf(c, d));
}

// log2(e): the blend works in log2 units, so the exponential is one v_exp_f32.
constexpr float kLog2e = 1.4426950408889634f;

// A splat's conic staged for the blend and its adjoint: log2(e) power = qa dx^2 + qb dx dy +
// qc dy^2 with qa = -log2(e) A / 2, qb = -log2(e) B, qc = -log2(e) C / 2.
__device__ inline float3 stagedConic(float A, float B, float C) {
    return make_float3(-0.5f * kLog2e * A, -kLog2e * B, -0.5f * kLog2e * C);
}

// log2 of a splat's opacity, staged beside its conic: alpha = min(0.99, exp2(e)) with
// e = log2(e) power + log2(o), so the opacity costs no multiply in the walk.
__device__ inline float stagedLogOpacity(float o) { return __builtin_amdgcn_logf(o); }

// e = log2(e) power + log2(o) at offset (dx, dy) from the centre, as
// (qa dx + qb dy) dx + ((qc dy) dy + lo) in three fused multiply-adds and two products. Every
// term of the power rounds relative to itself, so e > lo (the reference's power > 0 skip) holds
// only where the true power is zero to within rounding, as in the reference's own evaluation.
__device__ inline float splatExponent(float qa, float qb, float qc, float lo, float dx, float dy) {
    const float t = __builtin_fmaf(qa, dx, qb * dy);
    const float r = qc * dy;
    return __builtin_fmaf(t, dx, __builtin_fmaf(r, dy, lo));
}

// Front-to-back blend, one workgroup per TW x TH tile, one lane per pixel; each wavefront owns
// an 8 x 8 block of the tile. The per-pixel rule is the reference's (src/gpu_render.cu:180-233,
// src/render.cpp:193-221): skip power > 0; alpha = min(0.99, o exp(power)); skip
// alpha < 1/255; stop once T (1 - alpha) < 1e-4; C += rgb alpha T. Here power is evaluated in
// log2 units with fused multiply-adds and the opacity folded in (splatExponent), exp as exp2,
// and T (1 - alpha) as T - alpha T; these differ from the reference's rounding in the last bits,
// which moves alpha across 1/255 or T across 1e-4 only at the margins (at most 1/255 on a
// pixel).
//
// Per round, up to BATCH splat records of the tile's depth-sorted list are staged in LDS. With
// CULL, the staging lane also tests the splat against each wavefront's 8 x 8 block: a splat
// whose region alpha >= 1/255 (q <= 2 ln(255 o), widened by 5% + 0.1 for rounding) misses the
// block cannot change any of its pixels, so the block's wavefront never visits it. Each
// wavefront compacts the indices it must visit with ballots, walks them in order, and stops as
// soon as all of its pixels are done; the workgroup stops when all wavefronts are.
//
// A done pixel needs no flag of its own: its T makes every later splat a no-op. Without SAVE a
// stopping splat leaves T (1 - alpha) < 1e-4 in T (0 outside the image), so T (1 - alpha) of any
// later splat stays below 1e-4 and is never kept. With SAVE, where the final transmittance must
// survive, a stop stores -T (-1 outside the image): alpha T <= 0, T - alpha T < 1e-4, and |T| is
// the final transmittance.
//
// SAVE (frames the backward pass will read): pixState receives, per pixel, the final
// transmittance's bits and the position + 1 in the tile's list of the last splat blended into it
// (0: none).
template <int TW, int TH, int BATCH, bool CULL, bool SAVE>
__device__ inline void blendTile(int tileX, int tileY, uint2 range, const uint32_t* __restrict__ sortedGauss,
                                 const SplatRec* __restrict__ rec, int width, int height, uint8_t* __restrict__ out8,
                                 float* __restrict__ outf, uint2* __restrict__ pixState) {
    constexpr int kThreads = TW * TH;
    constexpr int kWaves = kThreads / kWave;
    constexpr int kWavesX = TW / 8;
    static_assert(kWaves <= 8, "at most eight 8x8 blocks per tile");
    // Staged per splat in three arrays of 16-byte rows read at one byte offset (the visit lists
    // hold offsets, so the walk computes no addresses): (u, v, qa, qb), (qc, log2 o, r, g), (b).
    __shared__ float4 sRec[3][BATCH];
    __shared__ uint8_t sMask[BATCH];
    __shared__ uint16_t sList[kWaves][BATCH + 4];  // padded: the walk reads up to 3 entries ahead
    auto row = [&](int r, uint32_t off) -> const float4& {
        return *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(sRec[r]) + off);
    };

    const int tid = threadIdx.x;
    const int wave = tid / kWave, lane = __lane_id();
    const int bx = tileX * TW + (wave % kWavesX) * 8, by = tileY * TH + (wave / kWavesX) * 8;
    const int x = bx + (lane & 7), y = by + (lane >> 3);
    const bool inside = x < width && y < height;
    const float px = x + 0.5f, py = y + 0.5f;
    const uint64_t below = lanesBelow();

    float T = inside ? 1.0f : (SAVE ? -1.0f : 0.0f);
    // Whether a pixel still blends (see the note on done pixels above).
    auto live = [](float t) { return SAVE ? t > 0.0f : t >= 1e-4f; };
    float cr = 0.0f, cg = 0.0f, cb = 0.0f;
    uint32_t last = 0;
    for (uint32_t start = range.x; start < range.y; start += BATCH) {
        if (__syncthreads_count(!live(T)) == kThreads) break;
        const int cnt = min(static_cast<uint32_t>(BATCH), range.y - start);
        for (int k = tid; k < cnt; k += kThreads) {
            const uint32_t g = sortedGauss[start + k];
            const float4 a = rec[g].geo;
            const float4 b = rec[g].geo2;
            const float4 col = rec[g].col;
            const float3 q = stagedConic(a.z, a.w, b.x);
            sRec[0][k] = make_float4(a.x, a.y, q.x, q.y);
            sRec[1][k] = make_float4(q.z, stagedLogOpacity(b.y), col.x, col.y);
            sRec[2][k].x = col.z;
            if constexpr (CULL) {
                uint32_t m = 0;
                const float op = b.y;
                if (op >= 1.0f / 255.0f) {
                    const float kq = fmaxf(2.0f * __logf(255.0f * op), 0.0f) * 1.05f + 0.1f;
                    const float A = a.z, B = a.w, C = b.x;
                    const float rB_C = -B / C, rB_A = -B / A;
#pragma unroll
                    for (int w = 0; w < kWaves; ++w) {
                        const float x0 = tileX * TW + (w % kWavesX) * 8 + 0.5f;
                        const float y0 = tileY * TH + (w / kWavesX) * 8 + 0.5f;
                        const float dx0 = x0 - a.x, dy0 = y0 - a.y;
                        if (quadMinOverBox(A, B, C, rB_C, rB_A, dx0, dx0 + 7.0f, dy0, dy0 + 7.0f) <= kq) m |= 1u << w;
                    }
                }
                sMask[k] = static_cast<uint8_t>(m);
            }
        }
        __syncthreads();
        if (__ballot(live(T)) == 0) continue;
        int count = cnt;
        if constexpr (CULL) {
            count = 0;
            for (int g0 = 0; g0 < cnt; g0 += kWave) {
                const int j = g0 + lane;
                const bool hit = j < cnt && ((sMask[j] >> wave) & 1u);
                const uint64_t bal = __ballot(hit);
                if (hit) sList[wave][count + __popcll(bal & below)] = static_cast<uint16_t>(j * 16);
                count += __popcll(bal);
            }
            waveSync();
        }
        // Software-pipelined walk over the list, two splats per step with two register sets:
        // while one splat is composited, the other set receives the next splat's records and the
        // list index after it is read. The per-pixel rule is applied with selects instead of
        // divergent branches; a skipped splat and a done pixel change nothing.
        auto composite = [&](const float4& a, const float4& b, float blue, uint32_t off) {
            const float e = splatExponent(a.z, a.w, b.x, b.y, px - a.x, py - a.y);
            const float alpha = fminf(0.99f, __builtin_amdgcn_exp2f(e));
            const float aT = alpha * T;
            const float nT = T - aT;
            const bool hit = !(e > b.y) && alpha >= 1.0f / 255.0f;
            const bool keep = nT >= 1e-4f;
            const bool use = hit && keep;
            // A weight of +0 leaves the (non-negative, finite) colour sums unchanged exactly.
            const float w = use ? aT : 0.0f;
            cr = __builtin_fmaf(b.z, w, cr);
            cg = __builtin_fmaf(b.w, w, cg);
            cb = __builtin_fmaf(blue, w, cb);
            if constexpr (SAVE) {
                const float stopped = keep ? nT : -__builtin_fabsf(T);
                T = hit ? stopped : T;
                last = use ? start - range.x + (off >> 4) + 1 : last;
            } else {
                T = hit ? nT : T;
            }
        };
        auto listAt = [&](int i) -> uint32_t { return CULL ? sList[wave][i] : static_cast<uint32_t>(i) * 16u; };
        if (count > 0) {
            uint32_t jA = listAt(0);
            float4 aA = row(0, jA), bA = row(1, jA);
            float cA = row(2, jA).x;
            uint32_t jB = listAt(1);  // entries past count are read, never used
            float4 aB, bB;
            float cB;
            for (int i = 0; i < count; i += 2) {
                const bool more1 = i + 1 < count;
                if (more1) {
                    aB = row(0, jB);
                    bB = row(1, jB);
                    cB = row(2, jB).x;
                }
                const uint32_t jNextA = listAt(i + 2);
                composite(aA, bA, cA, jA);
                if (!more1) break;
                jA = jNextA;
                if (i + 2 < count) {
                    aA = row(0, jA);
                    bA = row(1, jA);
                    cA = row(2, jA).x;
                }
                const uint32_t jNextB = listAt(i + 3);
                composite(aB, bB, cB, jB);
                jB = jNextB;
                if ((i & 7) == 6 && __ballot(live(T)) == 0) break;
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
        if constexpr (SAVE) pixState[p] = make_uint2(__float_as_uint(__builtin_fabsf(T)), last);
    }
}

// Workgroups take tiles in raster order: neighbouring tiles share splat records in L2, and a
// longest-list-first order measured slower on this path.
template <int TW, int TH, int BATCH, bool CULL, bool SAVE>
__global__ __launch_bounds__(TW* TH) void k_blend(const uint2* __restrict__ ranges,
                                                  const uint32_t* __restrict__ sortedGauss,
                                                  const SplatRec* __restrict__ rec, int width, int height, int tilesX,
                                                  uint8_t* __restrict__ out8, float* __restrict__ outf,
                                                  uint2* __restrict__ pixState, FrameCtl* __restrict__ ctl,
                                                  FrameStamps* __restrict__ stamps) {
    const int tile = blockIdx.x;
    if (stamps && tile == 0 && threadIdx.x == 0) stamps->t[2] = wall_clock64();
    if (!ctl->overflow)
        blendTile<TW, TH, BATCH, CULL, SAVE>(tile % tilesX, tile / tilesX, ranges[tile], sortedGauss, rec, width,
                                             height, out8, outf, pixState);
    if (stamps) {
        // The workgroup that ends last reads the clock for the kernel's end.
        __syncthreads();
        if (threadIdx.x == 0 && atomicAdd(&ctl->finished, 1u) == gridDim.x - 1) stamps->t[3] = wall_clock64();
    }
}

}  // namespace ans
