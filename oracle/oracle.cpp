// CPU oracle for the Gaussian-splat rasterizer. TEST INFRASTRUCTURE ONLY: tests, the smoke
// check and the CPU-baseline leg of the benchmark load this library as the checker; the
// product path never calls it.
//
// It restates the reference CPU renderer (src/render.cpp, src/sh.cpp, src/camera.cpp:12-32,
// src/ply_loader.cpp:124-139) in three modes:
//
//   ORC_MODE_REFERENCE  the reference's arithmetic exactly: FP64 world->camera transform
//                       (camera.cpp:12-17), float everywhere else, std::exp in the blend. With
//                       8x8 tiles it reproduces the compiled reference's RGB8 bit for bit.
//   ORC_MODE_KEYEXACT   the same maths with the world->camera transform in FP32, evaluated
//                       left to right with no fused multiply-add, which is what the HIP
//                       preprocess kernel does. Splat records, keys, tile ranges and per-tile
//                       lists from this mode are the bit-exact contract for the GPU.
//   ORC_MODE_F64        every operation in double (constants are the reference's float
//                       constants widened). Used as the FP64 forward for finite differences
//                       and as the base of the FP64 reference gradient.
//
// Build with -ffp-contract=off: every a*b + c below must round twice, like the kernel.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

namespace {

enum { ORC_MODE_REFERENCE = 0, ORC_MODE_KEYEXACT = 1, ORC_MODE_F64 = 2 };

using Clock = std::chrono::steady_clock;
double msSince(Clock::time_point t) { return std::chrono::duration<double, std::milli>(Clock::now() - t).count(); }

}  // namespace

extern "C" {

// Pinhole camera, COLMAP convention (camera.h:16-26): X_cam = R * X_world + t, R row-major.
struct orc_camera {
    double R[9];
    double t[3];
    double fx, fy, cx, cy;
    int width, height;
};

struct orc_stats {
    int64_t visible;
    int64_t pairs;
    double project_ms, sort_ms, blend_ms;
};

}  // extern "C"

namespace {

// Real spherical-harmonic constants (sh.cpp:10-14, gaussian.h:43); the values are the
// closed forms 1/(2 sqrt(pi)), sqrt(3/(4 pi)), ... rounded to float.
constexpr float kC0 = 0.28209479177387814f;
constexpr float kC1 = 0.4886025119029199f;
constexpr float kC2[5] = {1.0925484305920792f, -1.0925484305920792f, 0.31539156525252005f, -1.0925484305920792f,
                          0.5462742152960396f};
constexpr float kC3[7] = {-0.5900435899266435f, 2.890611442640554f, -0.4570457994644658f, 0.3731763325901154f,
                          -0.4570457994644658f, 1.445305721320277f, -0.5900435899266435f};

struct SceneView {
    int64_t n;
    int shDegree;  // degree stored in the scene
    const float* pos;      // [n][3]
    const float* scale;    // [n][3], activated (standard deviations)
    const float* rot;      // [n][4], w x y z, normalized
    const float* opacity;  // [n], activated
    const float* sh;       // [n][coeffs][3]
    int coeffs() const { return (shDegree + 1) * (shDegree + 1); }
};

template <class Real>
struct Splat {
    Real u = 0, v = 0;
    Real A = 0, B = 0, C = 0;  // conic [[A, B], [B, C]]
    Real opacity = 0;
    Real depth = 0;
    Real rgb[3] = {0, 0, 0};
    int radius = 0;
    int x0 = 0, y0 = 0, x1 = -1, y1 = -1;  // tile rect, inclusive
};

// Per-camera constants. The float ones are derived on the host exactly as the GPU library does
// (double -> float casts, camera centre in double then cast: camera.cpp:19-23;
// limX = 1.3f * (float)(0.5 * W / fx): render.cpp:58-59).
template <class Real>
struct CamConst {
    Real R[9], t[3], eye[3];
    Real fx, fy, cx, cy, limX, limY;
    double Rd[9], td[3];
    int W, H;
};

template <class Real>
CamConst<Real> camConst(const orc_camera& c) {
    CamConst<Real> k{};
    for (int i = 0; i < 9; ++i) k.R[i] = static_cast<Real>(static_cast<float>(c.R[i])), k.Rd[i] = c.R[i];
    for (int i = 0; i < 3; ++i) k.t[i] = static_cast<Real>(static_cast<float>(c.t[i])), k.td[i] = c.t[i];
    for (int j = 0; j < 3; ++j) {
        const double e = -(c.R[0 * 3 + j] * c.t[0] + c.R[1 * 3 + j] * c.t[1] + c.R[2 * 3 + j] * c.t[2]);
        k.eye[j] = std::is_same_v<Real, double> ? static_cast<Real>(e) : static_cast<Real>(static_cast<float>(e));
    }
    if constexpr (std::is_same_v<Real, double>) {
        for (int i = 0; i < 9; ++i) k.R[i] = c.R[i];
        for (int i = 0; i < 3; ++i) k.t[i] = c.t[i];
    }
    k.fx = static_cast<Real>(static_cast<float>(c.fx));
    k.fy = static_cast<Real>(static_cast<float>(c.fy));
    k.cx = static_cast<Real>(static_cast<float>(c.cx));
    k.cy = static_cast<Real>(static_cast<float>(c.cy));
    k.limX = static_cast<Real>(1.3f * static_cast<float>(0.5 * c.width / c.fx));
    k.limY = static_cast<Real>(1.3f * static_cast<float>(0.5 * c.height / c.fy));
    k.W = c.width;
    k.H = c.height;
    return k;
}

// Quaternion (w, x, y, z) to rotation matrix (render.cpp:31-36).
template <class Real>
void rotation(const Real q[4], Real M[3][3]) {
    const Real w = q[0], x = q[1], y = q[2], z = q[3];
    M[0][0] = 1 - 2 * (y * y + z * z); M[0][1] = 2 * (x * y - w * z);     M[0][2] = 2 * (x * z + w * y);
    M[1][0] = 2 * (x * y + w * z);     M[1][1] = 1 - 2 * (x * x + z * z); M[1][2] = 2 * (y * z - w * x);
    M[2][0] = 2 * (x * z - w * y);     M[2][1] = 2 * (y * z + w * x);     M[2][2] = 1 - 2 * (x * x + y * y);
}

// SH colour (sh.cpp:18-58): basis b[k] of the unit direction, rgb = 0.5 + sum b_k sh_k, clamp >= 0.
template <class Real>
void shColor(const float* sh, int degree, Real x, Real y, Real z, Real rgb[3]) {
    Real b[16];
    int count = 1;
    b[0] = kC0;
    if (degree >= 1) {
        b[1] = -Real(kC1) * y;
        b[2] = Real(kC1) * z;
        b[3] = -Real(kC1) * x;
        count = 4;
    }
    if (degree >= 2) {
        const Real xx = x * x, yy = y * y, zz = z * z;
        b[4] = Real(kC2[0]) * x * y;
        b[5] = Real(kC2[1]) * y * z;
        b[6] = Real(kC2[2]) * (2 * zz - xx - yy);
        b[7] = Real(kC2[3]) * x * z;
        b[8] = Real(kC2[4]) * (xx - yy);
        count = 9;
        if (degree >= 3) {
            b[9] = Real(kC3[0]) * y * (3 * xx - yy);
            b[10] = Real(kC3[1]) * x * y * z;
            b[11] = Real(kC3[2]) * y * (4 * zz - xx - yy);
            b[12] = Real(kC3[3]) * z * (2 * zz - 3 * xx - 3 * yy);
            b[13] = Real(kC3[4]) * x * (4 * zz - xx - yy);
            b[14] = Real(kC3[5]) * z * (xx - yy);
            b[15] = Real(kC3[6]) * x * (xx - 3 * yy);
            count = 16;
        }
    }
    rgb[0] = rgb[1] = rgb[2] = Real(0.5f);
    for (int k = 0; k < count; ++k)
        for (int c = 0; c < 3; ++c) rgb[c] += b[k] * Real(sh[k * 3 + c]);
    for (int c = 0; c < 3; ++c) rgb[c] = std::max(Real(0), rgb[c]);
}

// Projects Gaussian i (render.cpp:39-104). Returns false if culled. Deliberate deviations from
// the reference, shared with the kernel: the near test is !(z >= 0.2) so a NaN depth is culled,
// and a non-finite centre or a radius of 2^24 pixels or more is culled (the reference would cast
// such values to int, where large ones are undefined).
template <class Real>
bool makeSplat(const SceneView& sc, int64_t i, const CamConst<Real>& k, int mode, int shDegree, float splatScale,
               int tileW, int tileH, int tilesX, int tilesY, Splat<Real>& s) {
    const float* gp = sc.pos + 3 * i;
    Real px, py, pz;
    if (mode == ORC_MODE_REFERENCE) {
        // camera.cpp:12-17: FP64 transform, cast to float.
        const double x = gp[0], y = gp[1], z = gp[2];
        px = static_cast<float>(k.Rd[0] * x + k.Rd[1] * y + k.Rd[2] * z + k.td[0]);
        py = static_cast<float>(k.Rd[3] * x + k.Rd[4] * y + k.Rd[5] * z + k.td[1]);
        pz = static_cast<float>(k.Rd[6] * x + k.Rd[7] * y + k.Rd[8] * z + k.td[2]);
    } else {
        const Real x = gp[0], y = gp[1], z = gp[2];
        px = k.R[0] * x + k.R[1] * y + k.R[2] * z + k.t[0];
        py = k.R[3] * x + k.R[4] * y + k.R[5] * z + k.t[1];
        pz = k.R[6] * x + k.R[7] * y + k.R[8] * z + k.t[2];
    }
    if (!(pz >= Real(0.2f))) return false;

    Real q[4];
    for (int j = 0; j < 4; ++j) q[j] = sc.rot[4 * i + j];
    Real Rq[3][3];
    rotation(q, Rq);
    const Real sc3[3] = {Real(sc.scale[3 * i + 0] * splatScale), Real(sc.scale[3 * i + 1] * splatScale),
                         Real(sc.scale[3 * i + 2] * splatScale)};
    Real M[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) M[r][c] = Rq[r][c] * sc3[c];
    Real Sigma[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) Sigma[r][c] = M[r][0] * M[c][0] + M[r][1] * M[c][1] + M[r][2] * M[c][2];

    const Real tx = std::clamp(px / pz, -k.limX, k.limX) * pz;
    const Real ty = std::clamp(py / pz, -k.limY, k.limY) * pz;
    const Real fx = k.fx, fy = k.fy;
    const Real J[2][3] = {{fx / pz, 0, -fx * tx / (pz * pz)}, {0, fy / pz, -fy * ty / (pz * pz)}};
    Real T[2][3];
    for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 3; ++c) T[r][c] = J[r][0] * k.R[0 * 3 + c] + J[r][1] * k.R[1 * 3 + c] + J[r][2] * k.R[2 * 3 + c];
    Real TS[2][3];
    for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 3; ++c) TS[r][c] = T[r][0] * Sigma[0][c] + T[r][1] * Sigma[1][c] + T[r][2] * Sigma[2][c];
    const Real a = TS[0][0] * T[0][0] + TS[0][1] * T[0][1] + TS[0][2] * T[0][2] + Real(0.3f);
    const Real b = TS[0][0] * T[1][0] + TS[0][1] * T[1][1] + TS[0][2] * T[1][2];
    const Real c = TS[1][0] * T[1][0] + TS[1][1] * T[1][1] + TS[1][2] * T[1][2] + Real(0.3f);

    const Real det = a * c - b * b;
    if (!(det > 0)) return false;
    s.A = c / det;
    s.B = -b / det;
    s.C = a / det;

    const Real mid = Real(0.5f) * (a + c);
    const Real lambda = mid + std::sqrt(std::max(Real(0.1f), mid * mid - det));
    const Real rf = std::ceil(Real(3.0f) * std::sqrt(lambda));
    if (!(rf < Real(16777216.0f))) return false;
    s.radius = static_cast<int>(rf);

    s.u = fx * px / pz + k.cx;
    s.v = fy * py / pz + k.cy;
    if (!std::isfinite(s.u) || !std::isfinite(s.v)) return false;
    if (s.u + s.radius < 0 || s.v + s.radius < 0 || s.u - s.radius >= k.W || s.v - s.radius >= k.H) return false;

    // Tile rect (render.cpp:166-171; gpu_render.cu:133-134), truncating casts kept.
    s.x0 = std::max(0, static_cast<int>(s.u - s.radius) / tileW);
    s.y0 = std::max(0, static_cast<int>(s.v - s.radius) / tileH);
    s.x1 = std::min(tilesX - 1, static_cast<int>(s.u + s.radius) / tileW);
    s.y1 = std::min(tilesY - 1, static_cast<int>(s.v + s.radius) / tileH);

    s.depth = pz;
    s.opacity = sc.opacity[i];
    Real dir[3] = {Real(gp[0]) - k.eye[0], Real(gp[1]) - k.eye[1], Real(gp[2]) - k.eye[2]};
    const Real len = std::sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    shColor<Real>(sc.sh + static_cast<size_t>(i) * sc.coeffs() * 3, std::min(shDegree, sc.shDegree), dir[0] / len,
                  dir[1] / len, dir[2] / len, s.rgb);
    return true;
}

template <typename Fn>
void parallelFor(size_t n, unsigned threads, Fn fn) {
    if (threads == 0) threads = std::max(1u, std::thread::hardware_concurrency());
    if (threads == 1 || n <= 1) {
        for (size_t i = 0; i < n; ++i) fn(i);
        return;
    }
    std::atomic<size_t> next{0};
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < threads; ++t)
        pool.emplace_back([&] {
            for (size_t i; (i = next.fetch_add(1)) < n;) fn(i);
        });
    for (auto& th : pool) th.join();
}

// Order of visible Gaussians by depth, ties by ascending index (render.cpp:120-136: stable LSD
// radix sort on the float bits, three 11-bit passes over index-ascending input).
void sortByDepth(std::vector<uint32_t>& order, const std::vector<uint32_t>& depthBits) {
    const size_t n = order.size();
    std::vector<uint32_t> keys(n), keysTmp(n), idxTmp(n);
    for (size_t i = 0; i < n; ++i) keys[i] = depthBits[order[i]];
    for (int shift = 0; shift < 32; shift += 11) {
        std::vector<uint32_t> count(2049, 0);
        for (size_t i = 0; i < n; ++i) ++count[((keys[i] >> shift) & 2047) + 1];
        for (int b = 0; b < 2048; ++b) count[b + 1] += count[b];
        for (size_t i = 0; i < n; ++i) {
            const uint32_t dst = count[(keys[i] >> shift) & 2047]++;
            keysTmp[dst] = keys[i];
            idxTmp[dst] = order[i];
        }
        keys.swap(keysTmp);
        order.swap(idxTmp);
    }
}

}  // namespace

// A rendered frame with every intermediate the parity tests compare.
struct orc_frame {
    int mode = 0;
    int W = 0, H = 0, tileW = 8, tileH = 8, tilesX = 0, tilesY = 0;
    std::vector<Splat<float>> sf;
    std::vector<Splat<double>> sd;
    std::vector<char> alive;
    std::vector<uint32_t> order;      // visible Gaussians, depth-sorted
    std::vector<uint32_t> tileStart;  // tileCount + 1
    std::vector<uint32_t> binned;     // per-tile lists, concatenated
    std::vector<uint8_t> rgb;
    std::vector<double> rgbf;         // unclamped colour, FP64 copy of the float image
    std::vector<uint32_t> nContrib;   // per pixel: list entries consumed before stopping
    orc_stats st{};
};

namespace {

template <class Real>
std::vector<Splat<Real>>& splatsOf(orc_frame& f);
template <>
std::vector<Splat<float>>& splatsOf<float>(orc_frame& f) { return f.sf; }
template <>
std::vector<Splat<double>>& splatsOf<double>(orc_frame& f) { return f.sd; }

template <class Real>
void runFrame(orc_frame& f, const SceneView& sc, const orc_camera& cam, int shDegree, float splatScale,
              unsigned threads) {
    auto& splats = splatsOf<Real>(f);
    const CamConst<Real> k = camConst<Real>(cam);
    f.W = cam.width;
    f.H = cam.height;
    f.tilesX = (f.W + f.tileW - 1) / f.tileW;
    f.tilesY = (f.H + f.tileH - 1) / f.tileH;

    // 1. Project (render.cpp:144-157).
    auto t0 = Clock::now();
    const size_t n = static_cast<size_t>(sc.n);
    splats.assign(n, Splat<Real>{});
    f.alive.assign(n, 0);
    parallelFor((n + 4095) / 4096, threads, [&](size_t chunk) {
        const size_t end = std::min(n, (chunk + 1) * 4096);
        for (size_t i = chunk * 4096; i < end; ++i)
            f.alive[i] = makeSplat<Real>(sc, static_cast<int64_t>(i), k, f.mode, shDegree, splatScale, f.tileW, f.tileH,
                                         f.tilesX, f.tilesY, splats[i])
                             ? 1
                             : 0;
    });
    f.order.clear();
    for (size_t i = 0; i < n; ++i)
        if (f.alive[i]) f.order.push_back(static_cast<uint32_t>(i));
    f.st.visible = static_cast<int64_t>(f.order.size());
    f.st.project_ms = msSince(t0);

    // 2. Sort by depth, then bin (render.cpp:159-190).
    t0 = Clock::now();
    if constexpr (std::is_same_v<Real, float>) {
        std::vector<uint32_t> bits(n, 0);
        for (uint32_t i : f.order) std::memcpy(&bits[i], &splats[i].depth, 4);
        sortByDepth(f.order, bits);
    } else {
        std::stable_sort(f.order.begin(), f.order.end(),
                         [&](uint32_t a, uint32_t b) { return splats[a].depth < splats[b].depth; });
    }
    const size_t tileCount = static_cast<size_t>(f.tilesX) * f.tilesY;
    f.tileStart.assign(tileCount + 1, 0);
    for (uint32_t idx : f.order) {
        const auto& s = splats[idx];
        for (int ty = s.y0; ty <= s.y1; ++ty)
            for (int tx = s.x0; tx <= s.x1; ++tx) ++f.tileStart[static_cast<size_t>(ty) * f.tilesX + tx + 1];
    }
    for (size_t t = 0; t < tileCount; ++t) f.tileStart[t + 1] += f.tileStart[t];
    f.binned.assign(f.tileStart[tileCount], 0);
    std::vector<uint32_t> fill(f.tileStart.begin(), f.tileStart.end() - 1);
    for (uint32_t idx : f.order) {
        const auto& s = splats[idx];
        for (int ty = s.y0; ty <= s.y1; ++ty)
            for (int tx = s.x0; tx <= s.x1; ++tx) f.binned[fill[static_cast<size_t>(ty) * f.tilesX + tx]++] = idx;
    }
    f.st.pairs = static_cast<int64_t>(f.binned.size());
    f.st.sort_ms = msSince(t0);

    // 3. Blend front to back (render.cpp:192-221).
    t0 = Clock::now();
    f.rgb.assign(static_cast<size_t>(f.W) * f.H * 3, 0);
    f.rgbf.assign(static_cast<size_t>(f.W) * f.H * 3, 0.0);
    f.nContrib.assign(static_cast<size_t>(f.W) * f.H, 0);
    parallelFor(tileCount, threads, [&](size_t t) {
        const int tx = static_cast<int>(t % f.tilesX), ty = static_cast<int>(t / f.tilesX);
        const uint32_t* first = f.binned.data() + f.tileStart[t];
        const uint32_t* last = f.binned.data() + f.tileStart[t + 1];
        for (int y = ty * f.tileH; y < std::min(f.H, (ty + 1) * f.tileH); ++y) {
            for (int x = tx * f.tileW; x < std::min(f.W, (tx + 1) * f.tileW); ++x) {
                const Real px = Real(x + 0.5f), py = Real(y + 0.5f);
                Real T = 1, C[3] = {0, 0, 0};
                const uint32_t* it = first;
                for (; it != last; ++it) {
                    const auto& s = splats[*it];
                    const Real dx = px - s.u, dy = py - s.v;
                    const Real power = Real(-0.5f) * (s.A * dx * dx + s.C * dy * dy) - s.B * dx * dy;
                    if (power > 0) continue;
                    const Real alpha = std::min(Real(0.99f), s.opacity * std::exp(power));
                    if (alpha < Real(1.0f / 255.0f)) continue;
                    const Real nextT = T * (1 - alpha);
                    if (nextT < Real(1e-4f)) break;
                    for (int c = 0; c < 3; ++c) C[c] += s.rgb[c] * alpha * T;
                    T = nextT;
                }
                const size_t p = static_cast<size_t>(y) * f.W + x;
                f.nContrib[p] = static_cast<uint32_t>(it - first);
                for (int c = 0; c < 3; ++c) {
                    f.rgbf[p * 3 + c] = C[c];
                    f.rgb[p * 3 + c] = static_cast<uint8_t>(
                        static_cast<float>(std::clamp(C[c], Real(0), Real(1)) * Real(255.0f) + Real(0.5f)));
                }
            }
        }
    });
    f.st.blend_ms = msSince(t0);
}

}  // namespace

extern "C" {

const char* orc_version() { return "aniso-oracle-1"; }

// Renders one frame. sh is [n][(sh_degree_scene+1)^2][3]; rot is (w, x, y, z). Returns a frame
// handle the accessors below read; free it with orc_frame_free.
orc_frame* orc_frame_run(int mode, int64_t n, int scene_sh_degree, const float* pos, const float* scale,
                         const float* rot, const float* opacity, const float* sh, const orc_camera* cam, int sh_degree,
                         float splat_scale, int tile_w, int tile_h, int threads) {
    if (!cam || n < 0 || tile_w <= 0 || tile_h <= 0 || cam->width <= 0 || cam->height <= 0) return nullptr;
    auto f = std::make_unique<orc_frame>();
    f->mode = mode;
    f->tileW = tile_w;
    f->tileH = tile_h;
    const SceneView sc{n, scene_sh_degree, pos, scale, rot, opacity, sh};
    if (mode == ORC_MODE_F64)
        runFrame<double>(*f, sc, *cam, sh_degree, splat_scale, static_cast<unsigned>(threads));
    else
        runFrame<float>(*f, sc, *cam, sh_degree, splat_scale, static_cast<unsigned>(threads));
    return f.release();
}

void orc_frame_free(orc_frame* f) { delete f; }

void orc_frame_stats(const orc_frame* f, orc_stats* out) { *out = f->st; }

int64_t orc_frame_tile_count(const orc_frame* f) { return static_cast<int64_t>(f->tilesX) * f->tilesY; }

void orc_frame_image(const orc_frame* f, uint8_t* rgb, double* rgbf) {
    if (rgb) std::memcpy(rgb, f->rgb.data(), f->rgb.size());
    if (rgbf) std::memcpy(rgbf, f->rgbf.data(), f->rgbf.size() * sizeof(double));
}

void orc_frame_ncontrib(const orc_frame* f, uint32_t* out) {
    std::memcpy(out, f->nContrib.data(), f->nContrib.size() * 4);
}

// Sorted pairs as the GPU produces them: key = tile << 32 | float bits of depth, value = index;
// ranges[t] = [first, last) into the sorted list (gpu_render.cu:152-160, :164-177). Float modes.
void orc_frame_pairs(const orc_frame* f, uint64_t* keys, uint32_t* values, uint32_t* ranges) {
    const size_t tiles = static_cast<size_t>(f->tilesX) * f->tilesY;
    for (size_t t = 0; t < tiles; ++t) {
        if (ranges) {
            const bool empty = f->tileStart[t] == f->tileStart[t + 1];
            ranges[2 * t] = empty ? 0 : f->tileStart[t];
            ranges[2 * t + 1] = empty ? 0 : f->tileStart[t + 1];
        }
        for (uint32_t p = f->tileStart[t]; p < f->tileStart[t + 1]; ++p) {
            const uint32_t g = f->binned[p];
            if (keys) {
                uint32_t bits = 0;
                if (f->mode != ORC_MODE_F64) std::memcpy(&bits, &f->sf[g].depth, 4);
                keys[p] = (static_cast<uint64_t>(t) << 32) | bits;
            }
            if (values) values[p] = g;
        }
    }
}

// Per-Gaussian splat records: rec[i] = {u, v, A, B, C, opacity, depth, r, g, b} (culled: zeros),
// rect[i] = {x0, y0, x1, y1}, radius[i], alive[i].
void orc_frame_splats(const orc_frame* f, double* rec, int32_t* rect, int32_t* radius, uint8_t* alive) {
    const size_t n = f->alive.size();
    for (size_t i = 0; i < n; ++i) {
        auto put = [&](const auto& s) {
            if (rec) {
                double* r = rec + 10 * i;
                r[0] = s.u; r[1] = s.v; r[2] = s.A; r[3] = s.B; r[4] = s.C; r[5] = s.opacity; r[6] = s.depth;
                r[7] = s.rgb[0]; r[8] = s.rgb[1]; r[9] = s.rgb[2];
            }
            if (rect) {
                rect[4 * i + 0] = s.x0; rect[4 * i + 1] = s.y0; rect[4 * i + 2] = s.x1; rect[4 * i + 3] = s.y1;
            }
            if (radius) radius[i] = s.radius;
        };
        if (f->mode == ORC_MODE_F64) put(f->sd[i]); else put(f->sf[i]);
        if (alive) alive[i] = static_cast<uint8_t>(f->alive[i]);
    }
}

// Float bit patterns of the per-Gaussian records in a float mode, for bit-exact comparison with
// the kernel: out[i] = {u, v, A, B, C, opacity, depth, r, g, b} as uint32.
void orc_frame_splat_bits(const orc_frame* f, uint32_t* out) {
    const size_t n = f->sf.size();
    for (size_t i = 0; i < n; ++i) {
        const auto& s = f->sf[i];
        const float v[10] = {s.u, s.v, s.A, s.B, s.C, s.opacity, s.depth, s.rgb[0], s.rgb[1], s.rgb[2]};
        std::memcpy(out + 10 * i, v, sizeof v);
    }
}

// Activation of raw 3DGS values (ply_loader.cpp:124-139): scale = exp(log-scale),
// opacity = sigmoid(logit), rotation = q / |q| (identity if |q| = 0). Float arithmetic, std::exp.
void orc_activate(int64_t n, const float* log_scale, const float* logit, const float* quat, float* scale,
                  float* opacity, float* rot) {
    for (int64_t i = 0; i < n; ++i) {
        for (int j = 0; j < 3; ++j) scale[3 * i + j] = std::exp(log_scale[3 * i + j]);
        opacity[i] = 1.0f / (1.0f + std::exp(-logit[i]));
        const float* q = quat + 4 * i;
        const float nq = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
        for (int j = 0; j < 4; ++j) rot[4 * i + j] = nq > 0 ? q[j] / nq : (j == 0 ? 1.0f : 0.0f);
    }
}

}  // extern "C"

#include <random>

extern "C" {

// The random scene of the reference's GPU-vs-CPU test (tests/test_gpu_render.cpp:41-58):
// std::mt19937(seed), uniform_real_distribution<float>(-1, 1) and (0, 1), drawn in the same
// order; SH degree 3. Outputs activated arrays.
void orc_mt19937_scene(uint32_t seed, int64_t n, float* pos, float* scale, float* rot, float* opacity, float* sh) {
    std::mt19937 rng(seed);
    std::uniform_real_distribution<float> u(-1, 1), pos01(0, 1);
    for (int64_t i = 0; i < n; ++i) {
        // Braced initializers evaluate left to right.
        const float p[3] = {u(rng) * 2, u(rng) * 1.2f, 3 + pos01(rng) * 4};
        const float s[3] = {0.02f + 0.1f * pos01(rng), 0.02f + 0.05f * pos01(rng), 0.01f + 0.03f * pos01(rng)};
        const float q[4] = {u(rng), u(rng), u(rng), u(rng)};
        const float len = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
        for (int j = 0; j < 3; ++j) pos[3 * i + j] = p[j], scale[3 * i + j] = s[j];
        for (int j = 0; j < 4; ++j) rot[4 * i + j] = q[j] / len;
        opacity[i] = 0.2f + 0.8f * pos01(rng);
        for (int k = 0; k < 48; ++k) sh[48 * i + k] = k < 3 ? u(rng) * 1.5f : u(rng) * 0.15f;
    }
}

}  // extern "C"
