// Activation of raw 3DGS parameters, as the reference PLY loader performs it
// (src/ply_loader.cpp:124-130), in FP32 on the host or the device:
//   scale    = exp(log-scale)
//   opacity  = 1 / (1 + exp(-logit))
//   rotation = q / sqrt(w w + x x + y y + z z), or (1, 0, 0, 0) when that length is not positive
//
// The reference calls std::exp(float), which is glibc's expf on its platform. For the device to
// produce the same bits, expfTable restates the published algorithm of that function (the
// table-driven expf of Arm's optimized-routines, glibc's expf since 2.27): with N = 32,
//   z = x * N / ln 2 in double; k = z rounded to nearest (adding and subtracting 1.5 * 2^52);
//   r = z - k; 2^(k/N) = table[k mod N] with the exponent bits of floor(k/N) added;
//   2^(r/N) ~= 1 + c2 r + c1 r^2 + c0 r^3 (a minimax cubic, Horner in two halves);
//   the double product is rounded to float once.
// Out of range: x > 0x1.62e42ep6 overflows to +inf, x < -0x1.9fe368p6 underflows to +0, NaN
// propagates. The host test compares expfTable with the C library's expf on every 7th float.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ans {

// 2^(i/32) rounded to double, minus i << 47 (so adding k << 47 yields 2^(k/32)'s bits).
__host__ __device__ inline uint64_t exp2Table(int i) {
    constexpr uint64_t T[32] = {
        0x3ff0000000000000ull, 0x3fefd9b0d3158574ull, 0x3fefb5586cf9890full, 0x3fef9301d0125b51ull,
        0x3fef72b83c7d517bull, 0x3fef54873168b9aaull, 0x3fef387a6e756238ull, 0x3fef1e9df51fdee1ull,
        0x3fef06fe0a31b715ull, 0x3feef1a7373aa9cbull, 0x3feedea64c123422ull, 0x3feece086061892dull,
        0x3feebfdad5362a27ull, 0x3feeb42b569d4f82ull, 0x3feeab07dd485429ull, 0x3feea47eb03a5585ull,
        0x3feea09e667f3bcdull, 0x3fee9f75e8ec5f74ull, 0x3feea11473eb0187ull, 0x3feea589994cce13ull,
        0x3feeace5422aa0dbull, 0x3feeb737b0cdc5e5ull, 0x3feec49182a3f090ull, 0x3feed503b23e255dull,
        0x3feee89f995ad3adull, 0x3feeff76f2fb5e47ull, 0x3fef199bdd85529cull, 0x3fef3720dcef9069ull,
        0x3fef5818dcfba487ull, 0x3fef7c97337b9b5full, 0x3fefa4afa2a490daull, 0x3fefd0765b6e4540ull,
    };
    return T[i];
}

__host__ __device__ inline float expfTable(float x) {
    if (x != x) return x + x;
    if (x > 0x1.62e42ep6f) return __builtin_inff();
    if (x < -0x1.9fe368p6f) return 0.0f;
    constexpr double kInvLn2N = 0x1.71547652b82fep+0 * 32;
    constexpr double kShift = 0x1.8p+52;
    constexpr double c0 = 0x1.c6af84b912394p-5 / 32 / 32 / 32;
    constexpr double c1 = 0x1.ebfce50fac4f3p-3 / 32 / 32;
    constexpr double c2 = 0x1.62e42ff0c52d6p-1 / 32;
    const double z = kInvLn2N * static_cast<double>(x);
    double kd = z + kShift;
    const uint64_t ki = __builtin_bit_cast(uint64_t, kd);
    kd -= kShift;
    const double r = z - kd;
    const double s = __builtin_bit_cast(double, exp2Table(static_cast<int>(ki % 32)) + (ki << 47));
    const double p = c0 * r + c1;
    const double r2 = r * r;
    double y = c2 * r + 1.0;
    y = p * r2 + y;
    return static_cast<float>(y * s);
}

__host__ __device__ inline float activateScale(float logScale) { return expfTable(logScale); }

__host__ __device__ inline float activateOpacity(float logit) { return 1.0f / (1.0f + expfTable(-logit)); }

// q = (w, x, y, z) in, normalized out.
__host__ __device__ inline void activateRotation(const float q[4], float out[4]) {
    const float n = sqrtf(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    if (n > 0.0f) {
        for (int j = 0; j < 4; ++j) out[j] = q[j] / n;
    } else {
        out[0] = 1.0f;
        out[1] = out[2] = out[3] = 0.0f;
    }
}

}  // namespace ans
