// FP64 differentiable restatement of the rasterizer, for gradient checks. TEST INFRASTRUCTURE
// ONLY: tests load it as the checker; the product path never calls it.
//
// The forward is the reference CPU renderer's maths (src/render.cpp:39-104, :193-221,
// src/sh.cpp:18-58) in double, taking the raw 3DGS parameters and applying the loader's
// activations (src/ply_loader.cpp:124-139: scale = exp(log-scale), opacity = sigmoid(logit),
// rotation = q / |q|). Tiles are a parameter and per-tile lists may be supplied from outside
// (the GPU's), so the FP64 function is evaluated on exactly the GPU's list structure.
//
// orc_grad_run:      forward image + analytic adjoint for a given dL/dC (the image before the
//                    clamp and quantization), per Gaussian and raw parameter.
// orc_grad_fd:       central finite differences of L = sum(dLdC * C) for chosen Gaussians, each
//                    evaluated by re-blending only the tiles that Gaussian touches; a flag marks
//                    samples whose perturbation changed which splats contribute to a pixel, the
//                    0.99 clamps, the 1/255 skips, a termination point or a rect.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

constexpr double kC0 = 0.28209479177387814f;
constexpr double kC1 = 0.4886025119029199f;
constexpr double kC2[5] = {1.0925484305920792f, -1.0925484305920792f, 0.31539156525252005f, -1.0925484305920792f,
                           0.5462742152960396f};
constexpr double kC3[7] = {-0.5900435899266435f, 2.890611442640554f, -0.4570457994644658f, 0.3731763325901154f,
                           -0.4570457994644658f, 1.445305721320277f, -0.5900435899266435f};
constexpr int kParams = 59;  // 3 means, 3 log-scales, 4 quaternion, 1 logit, 48 SH

struct Cam {
    double R[9], t[3], eye[3];
    double fx, fy, cx, cy, limX, limY;
    int W, H, tileW, tileH, tilesX, tilesY;
};

struct Params {
    double m[3], ls[3], q[4], logit, sh[48];
};

// One projected Gaussian plus what its adjoint needs.
struct Proj {
    bool alive = false;
    double u, v, A, B, C, o, rgb[3];
    int x0, y0, x1, y1;
    // intermediates
    double s[3], qn[4], qlen, Rq[3][3], M[3][3], Sigma[3][3];
    double p[3], tx, ty, J[2][3], T[2][3], a, b, c, det;
    bool clampX, clampY;
    double dir[3], dlen, basis[16], rgbRaw[3];
    int degree;
};

void basisOf(double x, double y, double z, int degree, double* b) {
    for (int k = 0; k < 16; ++k) b[k] = 0;
    b[0] = kC0;
    if (degree >= 1) {
        b[1] = -kC1 * y;
        b[2] = kC1 * z;
        b[3] = -kC1 * x;
    }
    if (degree >= 2) {
        const double xx = x * x, yy = y * y, zz = z * z;
        b[4] = kC2[0] * x * y;
        b[5] = kC2[1] * y * z;
        b[6] = kC2[2] * (2 * zz - xx - yy);
        b[7] = kC2[3] * x * z;
        b[8] = kC2[4] * (xx - yy);
        if (degree >= 3) {
            b[9] = kC3[0] * y * (3 * xx - yy);
            b[10] = kC3[1] * x * y * z;
            b[11] = kC3[2] * y * (4 * zz - xx - yy);
            b[12] = kC3[3] * z * (2 * zz - 3 * xx - 3 * yy);
            b[13] = kC3[4] * x * (4 * zz - xx - yy);
            b[14] = kC3[5] * z * (xx - yy);
            b[15] = kC3[6] * x * (xx - 3 * yy);
        }
    }
}

// d basis_k / d(x, y, z) for the basis above (product rule on each polynomial).
void basisGrad(double x, double y, double z, int degree, double (*g)[3]) {
    for (int k = 0; k < 16; ++k) g[k][0] = g[k][1] = g[k][2] = 0;
    if (degree >= 1) {
        g[1][1] = -kC1;
        g[2][2] = kC1;
        g[3][0] = -kC1;
    }
    if (degree >= 2) {
        g[4][0] = kC2[0] * y; g[4][1] = kC2[0] * x;
        g[5][1] = kC2[1] * z; g[5][2] = kC2[1] * y;
        g[6][0] = kC2[2] * (-2 * x); g[6][1] = kC2[2] * (-2 * y); g[6][2] = kC2[2] * (4 * z);
        g[7][0] = kC2[3] * z; g[7][2] = kC2[3] * x;
        g[8][0] = kC2[4] * (2 * x); g[8][1] = kC2[4] * (-2 * y);
        if (degree >= 3) {
            const double xx = x * x, yy = y * y, zz = z * z;
            g[9][0] = kC3[0] * y * 6 * x; g[9][1] = kC3[0] * (3 * xx - 3 * yy);
            g[10][0] = kC3[1] * y * z; g[10][1] = kC3[1] * x * z; g[10][2] = kC3[1] * x * y;
            g[11][0] = kC3[2] * y * (-2 * x); g[11][1] = kC3[2] * (4 * zz - xx - 3 * yy); g[11][2] = kC3[2] * y * 8 * z;
            g[12][0] = kC3[3] * z * (-6 * x); g[12][1] = kC3[3] * z * (-6 * y);
            g[12][2] = kC3[3] * (6 * zz - 3 * xx - 3 * yy);
            g[13][0] = kC3[4] * (4 * zz - 3 * xx - yy); g[13][1] = kC3[4] * x * (-2 * y); g[13][2] = kC3[4] * x * 8 * z;
            g[14][0] = kC3[5] * z * 2 * x; g[14][1] = kC3[5] * z * (-2 * y); g[14][2] = kC3[5] * (xx - yy);
            g[15][0] = kC3[6] * (3 * xx - 3 * yy); g[15][1] = kC3[6] * x * (-6 * y);
        }
    }
}

bool project(const Params& P, const Cam& k, int shDegree, double splatScale, Proj& s) {
    s.alive = false;
    for (int r = 0; r < 3; ++r) s.p[r] = k.R[3 * r] * P.m[0] + k.R[3 * r + 1] * P.m[1] + k.R[3 * r + 2] * P.m[2] + k.t[r];
    const double px = s.p[0], py = s.p[1], pz = s.p[2];
    if (!(pz >= 0.2)) return false;
    for (int j = 0; j < 3; ++j) s.s[j] = std::exp(P.ls[j]) * splatScale;
    s.qlen = std::sqrt(P.q[0] * P.q[0] + P.q[1] * P.q[1] + P.q[2] * P.q[2] + P.q[3] * P.q[3]);
    for (int j = 0; j < 4; ++j) s.qn[j] = s.qlen > 0 ? P.q[j] / s.qlen : (j == 0 ? 1.0 : 0.0);
    const double w = s.qn[0], x = s.qn[1], y = s.qn[2], z = s.qn[3];
    double (&Rq)[3][3] = s.Rq;
    Rq[0][0] = 1 - 2 * (y * y + z * z); Rq[0][1] = 2 * (x * y - w * z);     Rq[0][2] = 2 * (x * z + w * y);
    Rq[1][0] = 2 * (x * y + w * z);     Rq[1][1] = 1 - 2 * (x * x + z * z); Rq[1][2] = 2 * (y * z - w * x);
    Rq[2][0] = 2 * (x * z - w * y);     Rq[2][1] = 2 * (y * z + w * x);     Rq[2][2] = 1 - 2 * (x * x + y * y);
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) s.M[r][c] = Rq[r][c] * s.s[c];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) s.Sigma[r][c] = s.M[r][0] * s.M[c][0] + s.M[r][1] * s.M[c][1] + s.M[r][2] * s.M[c][2];
    const double rx = px / pz, ry = py / pz;
    s.clampX = rx < -k.limX || rx > k.limX;
    s.clampY = ry < -k.limY || ry > k.limY;
    s.tx = std::clamp(rx, -k.limX, k.limX) * pz;
    s.ty = std::clamp(ry, -k.limY, k.limY) * pz;
    s.J[0][0] = k.fx / pz; s.J[0][1] = 0; s.J[0][2] = -k.fx * s.tx / (pz * pz);
    s.J[1][0] = 0; s.J[1][1] = k.fy / pz; s.J[1][2] = -k.fy * s.ty / (pz * pz);
    for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 3; ++c) s.T[r][c] = s.J[r][0] * k.R[c] + s.J[r][1] * k.R[3 + c] + s.J[r][2] * k.R[6 + c];
    double TS[2][3];
    for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 3; ++c) TS[r][c] = s.T[r][0] * s.Sigma[0][c] + s.T[r][1] * s.Sigma[1][c] + s.T[r][2] * s.Sigma[2][c];
    s.a = TS[0][0] * s.T[0][0] + TS[0][1] * s.T[0][1] + TS[0][2] * s.T[0][2] + double(0.3f);
    s.b = TS[0][0] * s.T[1][0] + TS[0][1] * s.T[1][1] + TS[0][2] * s.T[1][2];
    s.c = TS[1][0] * s.T[1][0] + TS[1][1] * s.T[1][1] + TS[1][2] * s.T[1][2] + double(0.3f);
    s.det = s.a * s.c - s.b * s.b;
    if (!(s.det > 0)) return false;
    s.A = s.c / s.det;
    s.B = -s.b / s.det;
    s.C = s.a / s.det;
    const double mid = 0.5 * (s.a + s.c);
    const double lambda = mid + std::sqrt(std::max(double(0.1f), mid * mid - s.det));
    const double rf = std::ceil(3.0 * std::sqrt(lambda));
    if (!(rf < 16777216.0)) return false;
    const int radius = static_cast<int>(rf);
    s.u = k.fx * px / pz + k.cx;
    s.v = k.fy * py / pz + k.cy;
    if (!std::isfinite(s.u) || !std::isfinite(s.v)) return false;
    if (s.u + radius < 0 || s.v + radius < 0 || s.u - radius >= k.W || s.v - radius >= k.H) return false;
    s.x0 = std::max(0, static_cast<int>(s.u - radius) / k.tileW);
    s.y0 = std::max(0, static_cast<int>(s.v - radius) / k.tileH);
    s.x1 = std::min(k.tilesX - 1, static_cast<int>(s.u + radius) / k.tileW);
    s.y1 = std::min(k.tilesY - 1, static_cast<int>(s.v + radius) / k.tileH);
    s.o = 1.0 / (1.0 + std::exp(-P.logit));
    for (int j = 0; j < 3; ++j) s.dir[j] = P.m[j] - k.eye[j];
    s.dlen = std::sqrt(s.dir[0] * s.dir[0] + s.dir[1] * s.dir[1] + s.dir[2] * s.dir[2]);
    for (int j = 0; j < 3; ++j) s.dir[j] /= s.dlen;
    s.degree = shDegree;
    basisOf(s.dir[0], s.dir[1], s.dir[2], shDegree, s.basis);
    const int count = (shDegree + 1) * (shDegree + 1);
    for (int ch = 0; ch < 3; ++ch) {
        double acc = 0.5;
        for (int kk = 0; kk < count; ++kk) acc += s.basis[kk] * P.sh[kk * 3 + ch];
        s.rgbRaw[ch] = acc;
        s.rgb[ch] = std::max(0.0, acc);
    }
    s.alive = true;
    return true;
}

// d(u, v, A, B, C, o, r, g, b) -> d(raw parameters) for one Gaussian (the adjoint of project()).
void projectAdjoint(const Params& P, const Cam& k, const Proj& s, const double* gs, double* out) {
    const double gu = gs[0], gv = gs[1], gA = gs[2], gB = gs[3], gC = gs[4], go = gs[5];
    for (int i = 0; i < kParams; ++i) out[i] = 0;
    const double px = s.p[0], py = s.p[1], pz = s.p[2];
    double gp[3] = {0, 0, 0};
    // u = fx px / pz + cx, v = fy py / pz + cy.
    gp[0] += gu * k.fx / pz;
    gp[1] += gv * k.fy / pz;
    gp[2] += -gu * k.fx * px / (pz * pz) - gv * k.fy * py / (pz * pz);
    // Conic of the 2D covariance: A = c / det, B = -b / det, C = a / det, det = a c - b^2.
    const double det2 = s.det * s.det;
    const double ga = gA * (-s.c * s.c / det2) + gB * (s.b * s.c / det2) + gC * (-s.b * s.b / det2);
    const double gb = gA * (2 * s.b * s.c / det2) + gB * (-1 / s.det - 2 * s.b * s.b / det2) + gC * (2 * s.a * s.b / det2);
    const double gc = gA * (-s.b * s.b / det2) + gB * (s.a * s.b / det2) + gC * (-s.a * s.a / det2);
    // cov = T Sigma T^T: a = cov00, b = cov01 = cov10, c = cov11.
    const double G[2][2] = {{ga, 0.5 * gb}, {0.5 * gb, gc}};
    double gSigma[3][3], gT[2][3];
    double TS[2][3];
    for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 3; ++c) TS[r][c] = s.T[r][0] * s.Sigma[0][c] + s.T[r][1] * s.Sigma[1][c] + s.T[r][2] * s.Sigma[2][c];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double acc = 0;
            for (int r = 0; r < 2; ++r)
                for (int c = 0; c < 2; ++c) acc += s.T[r][i] * G[r][c] * s.T[c][j];
            gSigma[i][j] = acc;
        }
    for (int r = 0; r < 2; ++r)
        for (int j = 0; j < 3; ++j) gT[r][j] = 2 * (G[r][0] * TS[0][j] + G[r][1] * TS[1][j]);
    // T = J W (W = camera rotation rows): dJ = dT W^T.
    double gJ[2][3];
    for (int r = 0; r < 2; ++r)
        for (int i = 0; i < 3; ++i) gJ[r][i] = gT[r][0] * k.R[3 * i] + gT[r][1] * k.R[3 * i + 1] + gT[r][2] * k.R[3 * i + 2];
    // J00 = fx/pz, J02 = -fx tx/pz^2, J11 = fy/pz, J12 = -fy ty/pz^2.
    const double pz2 = pz * pz, pz3 = pz2 * pz;
    gp[2] += gJ[0][0] * (-k.fx / pz2) + gJ[1][1] * (-k.fy / pz2) + gJ[0][2] * (2 * k.fx * s.tx / pz3) +
             gJ[1][2] * (2 * k.fy * s.ty / pz3);
    const double gtx = gJ[0][2] * (-k.fx / pz2), gty = gJ[1][2] * (-k.fy / pz2);
    // tx = clamp(px / pz, +-limX) pz.
    if (!s.clampX) gp[0] += gtx; else gp[2] += gtx * (s.tx / pz);
    if (!s.clampY) gp[1] += gty; else gp[2] += gty * (s.ty / pz);
    // p = R m + t.
    for (int j = 0; j < 3; ++j) out[j] += k.R[j] * gp[0] + k.R[3 + j] * gp[1] + k.R[6 + j] * gp[2];
    // Sigma = M M^T, M = Rq diag(s).
    double gM[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) gM[r][c] = 2 * (gSigma[r][0] * s.M[0][c] + gSigma[r][1] * s.M[1][c] + gSigma[r][2] * s.M[2][c]);
    double gRq[3][3], gsc[3] = {0, 0, 0};
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            gRq[r][c] = gM[r][c] * s.s[c];
            gsc[c] += gM[r][c] * s.Rq[r][c];
        }
    for (int j = 0; j < 3; ++j) out[3 + j] = gsc[j] * s.s[j];  // s = exp(ls) * splatScale
    // Rq(w, x, y, z).
    const double w = s.qn[0], x = s.qn[1], y = s.qn[2], z = s.qn[3];
    double gq[4];
    gq[0] = 2 * (-z * gRq[0][1] + y * gRq[0][2] + z * gRq[1][0] - x * gRq[1][2] - y * gRq[2][0] + x * gRq[2][1]);
    gq[1] = 2 * (y * gRq[0][1] + z * gRq[0][2] + y * gRq[1][0] - 2 * x * gRq[1][1] - w * gRq[1][2] + z * gRq[2][0] +
                 w * gRq[2][1] - 2 * x * gRq[2][2]);
    gq[2] = 2 * (-2 * y * gRq[0][0] + x * gRq[0][1] + w * gRq[0][2] + x * gRq[1][0] + z * gRq[1][2] - w * gRq[2][0] +
                 z * gRq[2][1] - 2 * y * gRq[2][2]);
    gq[3] = 2 * (-2 * z * gRq[0][0] - w * gRq[0][1] + x * gRq[0][2] + w * gRq[1][0] - 2 * z * gRq[1][1] + y * gRq[1][2] +
                 x * gRq[2][0] + y * gRq[2][1]);
    // q = q_raw / |q_raw|.
    if (s.qlen > 0) {
        const double dot = gq[0] * w + gq[1] * x + gq[2] * y + gq[3] * z;
        for (int j = 0; j < 4; ++j) out[6 + j] = (gq[j] - dot * s.qn[j]) / s.qlen;
    }
    // opacity = sigmoid(logit).
    out[10] = go * s.o * (1 - s.o);
    // SH colour, clamped at zero.
    double gdir[3] = {0, 0, 0};
    double gbasis[16] = {0};
    const int count = (s.degree + 1) * (s.degree + 1);
    for (int ch = 0; ch < 3; ++ch) {
        const double g = s.rgbRaw[ch] > 0 ? gs[6 + ch] : 0.0;
        for (int kk = 0; kk < count; ++kk) {
            out[11 + kk * 3 + ch] = g * s.basis[kk];
            gbasis[kk] += g * P.sh[kk * 3 + ch];
        }
    }
    double bg[16][3];
    basisGrad(s.dir[0], s.dir[1], s.dir[2], s.degree, bg);
    for (int kk = 0; kk < count; ++kk)
        for (int j = 0; j < 3; ++j) gdir[j] += gbasis[kk] * bg[kk][j];
    // dir = (m - eye) / |m - eye|.
    const double dd = gdir[0] * s.dir[0] + gdir[1] * s.dir[1] + gdir[2] * s.dir[2];
    for (int j = 0; j < 3; ++j) out[j] += (gdir[j] - dd * s.dir[j]) / s.dlen;
}

struct Scene64 {
    int64_t n;
    int shDegree;  // coefficients stored: 16 per channel required
    std::vector<Params> params;
};

Cam makeCam(const double* R, const double* t, double fx, double fy, double cx, double cy, int W, int H, int tw,
            int th) {
    Cam k{};
    for (int i = 0; i < 9; ++i) k.R[i] = R[i];
    for (int i = 0; i < 3; ++i) k.t[i] = t[i];
    for (int j = 0; j < 3; ++j) k.eye[j] = -(R[j] * t[0] + R[3 + j] * t[1] + R[6 + j] * t[2]);
    k.fx = fx; k.fy = fy; k.cx = cx; k.cy = cy;
    k.limX = double(1.3f) * (0.5 * W / fx);
    k.limY = double(1.3f) * (0.5 * H / fy);
    k.W = W; k.H = H; k.tileW = tw; k.tileH = th;
    k.tilesX = (W + tw - 1) / tw;
    k.tilesY = (H + th - 1) / th;
    return k;
}

template <typename Fn>
void parallelFor(size_t n, unsigned threads, Fn fn) {
    if (threads == 0) threads = std::max(1u, std::thread::hardware_concurrency());
    std::atomic<size_t> next{0};
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < threads; ++t)
        pool.emplace_back([&] {
            for (size_t i; (i = next.fetch_add(1)) < n;) fn(i);
        });
    for (auto& th : pool) th.join();
}

// Per-tile lists: by depth (ties by index), as the renderers bin them.
void binLists(const std::vector<Proj>& sp, const Cam& k, std::vector<uint32_t>& start, std::vector<uint32_t>& list) {
    std::vector<uint32_t> order;
    for (uint32_t i = 0; i < sp.size(); ++i)
        if (sp[i].alive) order.push_back(i);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return sp[a].p[2] < sp[b].p[2]; });
    const size_t tiles = static_cast<size_t>(k.tilesX) * k.tilesY;
    start.assign(tiles + 1, 0);
    for (uint32_t i : order)
        for (int ty = sp[i].y0; ty <= sp[i].y1; ++ty)
            for (int tx = sp[i].x0; tx <= sp[i].x1; ++tx) ++start[static_cast<size_t>(ty) * k.tilesX + tx + 1];
    for (size_t t = 0; t < tiles; ++t) start[t + 1] += start[t];
    list.assign(start[tiles], 0);
    std::vector<uint32_t> fill(start.begin(), start.end() - 1);
    for (uint32_t i : order)
        for (int ty = sp[i].y0; ty <= sp[i].y1; ++ty)
            for (int tx = sp[i].x0; tx <= sp[i].x1; ++tx) list[fill[static_cast<size_t>(ty) * k.tilesX + tx]++] = i;
}

Scene64 loadScene(int64_t n, int shDegree, const double* means, const double* logScale, const double* quat,
                  const double* logit, const double* sh) {
    Scene64 sc;
    sc.n = n;
    sc.shDegree = shDegree;
    sc.params.resize(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; ++i) {
        Params& p = sc.params[static_cast<size_t>(i)];
        for (int j = 0; j < 3; ++j) p.m[j] = means[3 * i + j], p.ls[j] = logScale[3 * i + j];
        for (int j = 0; j < 4; ++j) p.q[j] = quat[4 * i + j];
        p.logit = logit[i];
        for (int j = 0; j < 48; ++j) p.sh[j] = sh[48 * i + j];
    }
    return sc;
}

}  // namespace

extern "C" {

struct orc_grad_camera {
    double R[9];
    double t[3];
    double fx, fy, cx, cy;
    int width, height;
};

// Forward (image [H][W][3] of unclamped colour) and analytic adjoint (grad [n][59]: means,
// log-scales, quaternion (w, x, y, z), logit, SH [16][3]) for dLdC [H][W][3]. Scene SH must be
// degree-3 storage (48 values per Gaussian); sh_degree selects the evaluated degree.
// If list_start/list (tiles + 1, and the concatenated per-tile lists) are given, they replace the
// oracle's own binning. splat_grad (optional) receives [n][9] = dL/d{u, v, A, B, C, opacity, r,
// g, b}. Returns the number of visible Gaussians.
int64_t orc_grad_run(int64_t n, const double* means, const double* logScale, const double* quat, const double* logit,
                     const double* sh, const orc_grad_camera* cam, int sh_degree, double splat_scale, int tile_w,
                     int tile_h, const uint32_t* list_start, const uint32_t* list, const double* dLdC, double* image,
                     double* grad, double* splat_grad, int threads) {
    const Scene64 sc = loadScene(n, 3, means, logScale, quat, logit, sh);
    const Cam k = makeCam(cam->R, cam->t, cam->fx, cam->fy, cam->cx, cam->cy, cam->width, cam->height, tile_w, tile_h);
    std::vector<Proj> sp(static_cast<size_t>(n));
    parallelFor(static_cast<size_t>(n), threads, [&](size_t i) { project(sc.params[i], k, sh_degree, splat_scale, sp[i]); });
    int64_t visible = 0;
    for (const auto& s : sp) visible += s.alive;
    std::vector<uint32_t> start, lists;
    const size_t tiles = static_cast<size_t>(k.tilesX) * k.tilesY;
    if (list_start && list) {
        start.assign(list_start, list_start + tiles + 1);
        lists.assign(list, list + start[tiles]);
    } else {
        binLists(sp, k, start, lists);
    }
    // Per pair: d(u, v, A, B, C, o, r, g, b), reduced per Gaussian afterwards (deterministic).
    std::vector<double> partial(lists.size() * 9, 0.0);
    parallelFor(tiles, threads, [&](size_t t) {
        const int tx = static_cast<int>(t % k.tilesX), ty = static_cast<int>(t / k.tilesX);
        const uint32_t* first = lists.data() + start[t];
        const size_t len = start[t + 1] - start[t];
        std::vector<double> Ts, alphas;
        std::vector<size_t> idx;
        for (int y = ty * k.tileH; y < std::min(k.H, (ty + 1) * k.tileH); ++y)
            for (int x = tx * k.tileW; x < std::min(k.W, (tx + 1) * k.tileW); ++x) {
                const double px = x + 0.5, py = y + 0.5;
                const size_t pix = static_cast<size_t>(y) * k.W + x;
                // Forward, remembering each contributor's T and alpha.
                double T = 1.0, C[3] = {0, 0, 0};
                Ts.clear(); alphas.clear(); idx.clear();
                for (size_t i = 0; i < len; ++i) {
                    const Proj& s = sp[first[i]];
                    const double dx = px - s.u, dy = py - s.v;
                    const double power = -0.5 * (s.A * dx * dx + s.C * dy * dy) - s.B * dx * dy;
                    if (power > 0) continue;
                    const double alpha = std::min(double(0.99f), s.o * std::exp(power));
                    if (alpha < double(1.0f / 255.0f)) continue;
                    const double nextT = T * (1 - alpha);
                    if (nextT < double(1e-4f)) break;
                    for (int c = 0; c < 3; ++c) C[c] += s.rgb[c] * alpha * T;
                    Ts.push_back(T); alphas.push_back(alpha); idx.push_back(i);
                    T = nextT;
                }
                for (int c = 0; c < 3; ++c) image[3 * pix + c] = C[c];
                const double* gC = dLdC + 3 * pix;
                // Reverse: accum = colour blended behind splat i, per unit of T after it.
                double accum[3] = {0, 0, 0};
                for (size_t q = idx.size(); q-- > 0;) {
                    const size_t i = idx[q];
                    const Proj& s = sp[first[i]];
                    const double a = alphas[q], Ti = Ts[q];
                    double* pg = partial.data() + 9 * (start[t] + i);
                    double gAlpha = 0;
                    for (int c = 0; c < 3; ++c) {
                        pg[6 + c] += gC[c] * a * Ti;
                        gAlpha += gC[c] * Ti * (s.rgb[c] - accum[c]);
                        accum[c] = a * s.rgb[c] + (1 - a) * accum[c];
                    }
                    const double dx = px - s.u, dy = py - s.v;
                    const double power = -0.5 * (s.A * dx * dx + s.C * dy * dy) - s.B * dx * dy;
                    const double G = std::exp(power);
                    if (s.o * G > double(0.99f)) continue;  // clamped: alpha does not move
                    pg[5] += gAlpha * G;
                    const double gPow = gAlpha * s.o * G;
                    pg[0] += gPow * (s.A * dx + s.B * dy);
                    pg[1] += gPow * (s.C * dy + s.B * dx);
                    pg[2] += gPow * (-0.5 * dx * dx);
                    pg[3] += gPow * (-dx * dy);
                    pg[4] += gPow * (-0.5 * dy * dy);
                }
            }
    });
    std::vector<double> gs(static_cast<size_t>(n) * 9, 0.0);
    for (size_t p = 0; p < lists.size(); ++p)
        for (int j = 0; j < 9; ++j) gs[9 * lists[p] + j] += partial[9 * p + j];
    if (splat_grad) std::memcpy(splat_grad, gs.data(), gs.size() * sizeof(double));
    parallelFor(static_cast<size_t>(n), threads, [&](size_t i) {
        double* out = grad + kParams * i;
        if (!sp[i].alive) {
            for (int j = 0; j < kParams; ++j) out[j] = 0;
            return;
        }
        projectAdjoint(sc.params[i], k, sp[i], gs.data() + 9 * i, out);
    });
    return visible;
}

// Central differences of L = sum(dLdC * C) for Gaussians idx[0..m) and all 59 raw parameters,
// step h_rel * max(1, |x|) for parameter x. Each evaluation re-projects only the perturbed
// Gaussian and re-blends only the tiles in the union of its base and perturbed rects, with its
// list position re-derived from its new depth. flags[m][59]: 1 if either perturbation changed the
// structure signature of any re-blended pixel or the Gaussian's rect or depth order, 2 if it is
// culled in the base frame.
void orc_grad_fd(int64_t n, const double* means, const double* logScale, const double* quat, const double* logit,
                 const double* sh, const orc_grad_camera* cam, int sh_degree, double splat_scale, int tile_w,
                 int tile_h, const double* dLdC, const int64_t* idx, int64_t m, double h_rel, double* fd,
                 uint8_t* flags, int threads) {
    const Scene64 sc = loadScene(n, 3, means, logScale, quat, logit, sh);
    const Cam k = makeCam(cam->R, cam->t, cam->fx, cam->fy, cam->cx, cam->cy, cam->width, cam->height, tile_w, tile_h);
    std::vector<Proj> sp(static_cast<size_t>(n));
    parallelFor(static_cast<size_t>(n), threads, [&](size_t i) { project(sc.params[i], k, sh_degree, splat_scale, sp[i]); });
    std::vector<uint32_t> start, lists;
    binLists(sp, k, start, lists);
    auto depthLess = [&](uint32_t a, double da, uint32_t b, double db) { return da < db || (da == db && a < b); };

    parallelFor(static_cast<size_t>(m), threads, [&](size_t s) {
        const uint32_t g = static_cast<uint32_t>(idx[s]);
        uint8_t* fl = flags + kParams * s;
        double* out = fd + kParams * s;
        if (!sp[g].alive) {
            for (int j = 0; j < kParams; ++j) out[j] = 0, fl[j] = 2;
            return;
        }
        std::vector<Proj> local = {sp[g]};
        std::vector<uint32_t> tl;
        for (int j = 0; j < kParams; ++j) {
            Params P = sc.params[g];
            double* x = j < 3 ? &P.m[j] : j < 6 ? &P.ls[j - 3] : j < 10 ? &P.q[j - 6] : j == 10 ? &P.logit : &P.sh[j - 11];
            const double x0 = *x, h = h_rel * std::max(1.0, std::fabs(x0));
            double L[2] = {0, 0};
            bool changed = false;
            Proj pr[2];
            for (int side = 0; side < 2; ++side) {
                *x = x0 + (side ? -h : h);
                project(P, k, sh_degree, splat_scale, pr[side]);
                if (!pr[side].alive || pr[side].x0 != sp[g].x0 || pr[side].y0 != sp[g].y0 || pr[side].x1 != sp[g].x1 ||
                    pr[side].y1 != sp[g].y1)
                    changed = true;
            }
            *x = x0;
            if (changed) {
                out[j] = 0;
                fl[j] = 1;
                continue;
            }
            // Re-blend the tiles of g's rect with g replaced by its perturbed projection; its
            // list position is re-derived from its new depth and must not move.
            std::vector<uint8_t> sigs[3];
            for (int side = -1; side < 2 && !changed; ++side) {
                const Proj& sub = side < 0 ? sp[g] : pr[side];
                std::vector<uint8_t>& sg = sigs[side + 1];
                double sum = 0;
                for (int ty = sp[g].y0; ty <= sp[g].y1 && !changed; ++ty)
                    for (int tx = sp[g].x0; tx <= sp[g].x1 && !changed; ++tx) {
                        const size_t t = static_cast<size_t>(ty) * k.tilesX + tx;
                        tl.clear();
                        for (uint32_t p = start[t]; p < start[t + 1]; ++p)
                            if (lists[p] != g) tl.push_back(lists[p]);
                        size_t pos = 0, basePos = 0;
                        while (pos < tl.size() && depthLess(tl[pos], sp[tl[pos]].p[2], g, sub.p[2])) ++pos;
                        while (basePos < tl.size() && depthLess(tl[basePos], sp[tl[basePos]].p[2], g, sp[g].p[2]))
                            ++basePos;
                        if (pos != basePos) {
                            changed = true;
                            break;
                        }
                        tl.insert(tl.begin() + static_cast<long>(pos), g);
                        for (int y = ty * k.tileH; y < std::min(k.H, (ty + 1) * k.tileH); ++y)
                            for (int xq = tx * k.tileW; xq < std::min(k.W, (tx + 1) * k.tileW); ++xq) {
                                const size_t pix = static_cast<size_t>(y) * k.W + xq;
                                double T = 1.0, C[3] = {0, 0, 0};
                                for (size_t i = 0; i < tl.size(); ++i) {
                                    const Proj& q = tl[i] == g ? sub : sp[tl[i]];
                                    const double dx = xq + 0.5 - q.u, dy = y + 0.5 - q.v;
                                    const double power = -0.5 * (q.A * dx * dx + q.C * dy * dy) - q.B * dx * dy;
                                    if (power > 0) { sg.push_back(0); continue; }
                                    const double raw = q.o * std::exp(power);
                                    const double alpha = std::min(double(0.99f), raw);
                                    if (alpha < double(1.0f / 255.0f)) { sg.push_back(0); continue; }
                                    const double nextT = T * (1 - alpha);
                                    if (nextT < double(1e-4f)) { sg.push_back(3); break; }
                                    for (int c = 0; c < 3; ++c) C[c] += q.rgb[c] * alpha * T;
                                    T = nextT;
                                    sg.push_back(raw > double(0.99f) ? 2 : 1);
                                }
                                sg.push_back(9);
                                for (int c = 0; c < 3; ++c) sum += dLdC[3 * pix + c] * C[c];
                            }
                    }
                if (side >= 0) L[side] = sum;
            }
            if (!changed && (sigs[0] != sigs[1] || sigs[0] != sigs[2])) changed = true;
            out[j] = changed ? 0 : (L[0] - L[1]) / (2 * h);
            fl[j] = changed ? 1 : 0;
        }
    });
}

}  // extern "C"

extern "C" {

// The projection adjoint alone: grad [n][59] from per-Gaussian splat-record gradients
// splat_grad [n][9] = dL/d{u, v, A, B, C, opacity, r, g, b}.
void orc_grad_project_adjoint(int64_t n, const double* means, const double* logScale, const double* quat,
                              const double* logit, const double* sh, const orc_grad_camera* cam, int sh_degree,
                              double splat_scale, int tile_w, int tile_h, const double* splat_grad, double* grad) {
    const Scene64 sc = loadScene(n, 3, means, logScale, quat, logit, sh);
    const Cam k = makeCam(cam->R, cam->t, cam->fx, cam->fy, cam->cx, cam->cy, cam->width, cam->height, tile_w, tile_h);
    for (int64_t i = 0; i < n; ++i) {
        Proj s;
        double* out = grad + kParams * i;
        if (!project(sc.params[static_cast<size_t>(i)], k, sh_degree, splat_scale, s)) {
            for (int j = 0; j < kParams; ++j) out[j] = 0;
            continue;
        }
        projectAdjoint(sc.params[static_cast<size_t>(i)], k, s, splat_grad + 9 * i, out);
    }
}

}  // extern "C"
