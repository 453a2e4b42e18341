// Scene ingest on the device: 3DGS PLY files in and out.
//
// In: the reference loads a PLY on the CPU (src/ply_loader.cpp:77-141: parse, activate, reorder
// f_rest to [coefficient][channel]) and its GpuRenderer uploads the activated arrays
// (src/gpu_render.cu:267-292). Here the host parses only the header (ply.hpp); the vertex block
// goes to the device as it is, streamed from the file through two pinned buffers so the read of
// one chunk overlaps the copy of the previous one, and one kernel activates, drops the normals
// and writes the renderer's layout (positions, scales, unit quaternions, opacities, planar SH).
// The activation is bit-identical to the reference loader's (activate.hpp).
//
// Out: the training pipeline writes its parameters as a PLY (pipeline/train_gsplat.py:109-128);
// ans_ply_write builds the same file image from the differentiable path's device arrays, the
// interleave done by a kernel.
#include "activate.hpp"
#include "aniso_amd.h"
#include "engine.hpp"
#include "ply.hpp"

#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace ans {

namespace {

constexpr int kRows = 64;        // vertex rows per workgroup (one wavefront)
constexpr int kMaxStride = 256;  // properties per vertex the LDS staging takes (64 KB)
constexpr int kMaxRest = 72;     // f_rest columns of a degree-4 file

// Column indices of one file, passed by value to the kernel.
struct IngestCols {
    int stride;          // floats per vertex row
    int coeffs;          // SH coefficients per channel kept: (min(degree, 3) + 1)^2
    int restPerChannel;  // f_rest columns per channel in the file
    int pos[3], opacity, scale[3], rot[4], dc[3];
    int rest[kMaxRest];
};

// Copies `count` rows of `stride` floats starting at row `row0` into LDS: the rows are contiguous
// and the block's first row starts at a multiple of 64 * stride floats, so 16-byte loads.
__device__ inline void stageRows(const float* __restrict__ src, size_t base, int floats, float* rows) {
    const float4* src4 = reinterpret_cast<const float4*>(src + base);
    const int vec = floats / 4;
    for (int k = threadIdx.x; k < vec; k += kRows) reinterpret_cast<float4*>(rows)[k] = src4[k];
    for (int k = vec * 4 + static_cast<int>(threadIdx.x); k < floats; k += kRows) rows[k] = src[base + k];
}

// One wavefront per 64 vertex rows: stage, then each lane de-interleaves and activates its row
// (src/ply_loader.cpp:118-137): position; scale = exp; opacity = sigmoid; rotation normalized;
// SH coefficient 0 from f_dc, coefficient k + 1 of channel c from f_rest[c * restPerChannel + k].
// SH out is planar, [coefficient * 3 + channel][n].
__global__ __launch_bounds__(kRows) void k_plyIngest(const float* __restrict__ body, int64_t n, IngestCols c,
                                                     float* __restrict__ pos, float* __restrict__ scale,
                                                     float4* __restrict__ rot, float* __restrict__ opacity,
                                                     float* __restrict__ sh) {
    extern __shared__ float rows[];
    const int64_t row0 = static_cast<int64_t>(blockIdx.x) * kRows;
    const int count = static_cast<int>(min<int64_t>(kRows, n - row0));
    stageRows(body, static_cast<size_t>(row0) * c.stride, count * c.stride, rows);
    __syncthreads();
    const int t = threadIdx.x;
    if (t >= count) return;
    const float* v = rows + t * c.stride;
    const int64_t i = row0 + t;
    for (int k = 0; k < 3; ++k) pos[3 * i + k] = v[c.pos[k]];
    for (int k = 0; k < 3; ++k) scale[3 * i + k] = activateScale(v[c.scale[k]]);
    opacity[i] = activateOpacity(v[c.opacity]);
    const float q[4] = {v[c.rot[0]], v[c.rot[1]], v[c.rot[2]], v[c.rot[3]]};
    float u[4];
    activateRotation(q, u);
    rot[i] = make_float4(u[0], u[1], u[2], u[3]);
    for (int ch = 0; ch < 3; ++ch) sh[static_cast<size_t>(ch) * n + i] = v[c.dc[ch]];
#pragma unroll
    for (int k = 0; k < 15; ++k) {
        if (k + 1 >= c.coeffs) break;
        for (int ch = 0; ch < 3; ++ch)
            sh[static_cast<size_t>((k + 1) * 3 + ch) * n + i] = v[c.rest[ch * c.restPerChannel + k]];
    }
}

// The PLY row of the training pipeline's writer (pipeline/train_gsplat.py:109-123) for each
// Gaussian: x y z, nx ny nz = 0, f_dc_0..2, f_rest channel-major (all coefficients 1.. of red,
// then green, then blue), opacity logit, log-scales, the quaternion divided by max(|q|, 1e-12).
// Rows are assembled in LDS and stored as contiguous 16-byte writes.
__global__ __launch_bounds__(kRows) void k_plyRows(int64_t n, int coeffs, const float* __restrict__ means,
                                                   const float* __restrict__ logScales,
                                                   const float* __restrict__ quats,
                                                   const float* __restrict__ logits, const float* __restrict__ sh,
                                                   float* __restrict__ out) {
    extern __shared__ float rows[];
    const int stride = 17 + 3 * (coeffs - 1);
    const int64_t row0 = static_cast<int64_t>(blockIdx.x) * kRows;
    const int count = static_cast<int>(min<int64_t>(kRows, n - row0));
    const int t = threadIdx.x;
    if (t < count) {
        const int64_t i = row0 + t;
        float* r = rows + t * stride;
        for (int k = 0; k < 3; ++k) r[k] = means[3 * i + k];
        for (int k = 3; k < 6; ++k) r[k] = 0.0f;
        const float* s = sh + static_cast<size_t>(i) * coeffs * 3;
        for (int ch = 0; ch < 3; ++ch) r[6 + ch] = s[ch];
        const int rest = coeffs - 1;
        for (int ch = 0; ch < 3; ++ch)
            for (int k = 0; k < rest; ++k) r[9 + ch * rest + k] = s[(k + 1) * 3 + ch];
        int o = 9 + 3 * rest;
        r[o++] = logits[i];
        for (int k = 0; k < 3; ++k) r[o++] = logScales[3 * i + k];
        const float* q = quats + 4 * i;
        const float len = sqrtf(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
        const float d = len < 1e-12f ? 1e-12f : len;  // NaN stays NaN, as torch's clamp_min keeps it
        for (int k = 0; k < 4; ++k) r[o++] = q[k] / d;
    }
    __syncthreads();
    const int floats = count * stride;
    float* dst = out + static_cast<size_t>(row0) * stride;
    const int vec = floats / 4;
    for (int k = t; k < vec; k += kRows) reinterpret_cast<float4*>(dst)[k] = reinterpret_cast<const float4*>(rows)[k];
    for (int k = vec * 4 + t; k < floats; k += kRows) dst[k] = rows[k];
}

IngestCols columnsOf(const PlyLayout& L) {
    if (L.stride > kMaxStride)
        throw PlyError("PLY has " + std::to_string(L.stride) + " properties per vertex; at most " +
                       std::to_string(kMaxStride) + " are supported");
    IngestCols c{};
    c.stride = L.stride;
    const int kept = std::min(L.degree, 3);
    c.coeffs = (kept + 1) * (kept + 1);
    c.restPerChannel = (L.degree + 1) * (L.degree + 1) - 1;
    for (int k = 0; k < 3; ++k) c.pos[k] = L.pos[k], c.scale[k] = L.scale[k], c.dc[k] = L.dc[k];
    for (int k = 0; k < 4; ++k) c.rot[k] = L.rot[k];
    c.opacity = L.opacity;
    for (size_t k = 0; k < L.rest.size(); ++k) c.rest[k] = L.rest[k];
    return c;
}

// Runs the ingest kernel over the device copy of the vertex block into a new renderer.
ans_renderer* ingestBody(const PlyLayout& L, const float* devBody, std::unique_ptr<ans_renderer> r) {
    const IngestCols c = columnsOf(L);
    const int64_t n = L.n;
    const int kept = std::min(L.degree, 3);
    hipStream_t st = r->engine->stream;
    if (n > 0) {
        const uint32_t blocks = static_cast<uint32_t>((n + kRows - 1) / kRows);
        k_plyIngest<<<blocks, kRows, sizeof(float) * kRows * L.stride, st>>>(devBody, n, c, r->pos.p, r->scale.p,
                                                                            r->rot.p, r->opacity.p, r->sh.p);
        hipCheck(hipGetLastError(), "k_plyIngest launch");
    }
    hipCheck(hipStreamSynchronize(st), "hipStreamSynchronize");
    r->scene = SceneDev{n, kept, r->pos.p, r->scale.p, r->rot.p, r->opacity.p, r->sh.p, true, false};
    return r.release();
}

std::unique_ptr<ans_renderer> newRenderer(const PlyLayout& L) {
    if (L.n > (int64_t(1) << 30)) throw std::invalid_argument("Gaussian count must be in 0..2^30");
    int devices = 0;
    if (hipGetDeviceCount(&devices) != hipSuccess || devices == 0) throw std::runtime_error("no HIP device");
    auto r = std::make_unique<ans_renderer>();
    const size_t m = static_cast<size_t>(std::max<int64_t>(L.n, 1));
    const int kept = std::min(L.degree, 3);
    const size_t coeffs = static_cast<size_t>((kept + 1) * (kept + 1));
    r->pos.reserve(3 * m);
    r->scale.reserve(3 * m);
    r->rot.reserve(m);
    r->opacity.reserve(m);
    r->sh.reserve(coeffs * 3 * m);
    r->engine = std::make_unique<Engine>();
    return r;
}

// Process-wide pinned staging for file reads: two buffers, each with an event marking its last
// copy. Allocated on the first load and kept until the process ends (pinning memory costs more
// than a load of a typical scene); loads that share it are serialized by stagingMutex.
struct Staging {
    static constexpr size_t kBytes = size_t(16) << 20;
    uint8_t* buf[2] = {};
    hipEvent_t done[2] = {};
    Staging() {
        for (int k = 0; k < 2; ++k) {
            hipCheck(hipHostMalloc(reinterpret_cast<void**>(&buf[k]), kBytes, hipHostMallocDefault), "hipHostMalloc");
            hipCheck(hipEventCreate(&done[k]), "hipEventCreate");
        }
    }
};

std::mutex stagingMutex;

Staging& staging() {
    static Staging* s = new Staging();  // never freed: lives as long as the process
    return *s;
}

struct Fd {
    int fd = -1;
    ~Fd() {
        if (fd >= 0) ::close(fd);
    }
};

// Reads bytes at `offset` into dst with `threads` concurrent preads (the page-cache copy of one
// thread is slower than the copy to the device). False if the file ends early.
bool readAt(int fd, uint8_t* dst, size_t bytes, uint64_t offset, int threads) {
    std::atomic<bool> ok{true};
    const size_t part = (bytes + threads - 1) / threads;
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t) {
        const size_t b0 = std::min(bytes, t * part), b1 = std::min(bytes, b0 + part);
        pool.emplace_back([&, b0, b1] {
            for (size_t b = b0; b < b1;) {
                const ssize_t got = ::pread(fd, dst + b, b1 - b, static_cast<off_t>(offset + b));
                if (got <= 0) {
                    ok = false;
                    return;
                }
                b += static_cast<size_t>(got);
            }
        });
    }
    for (auto& th : pool) th.join();
    return ok;
}

ans_renderer* loadPly(const char* path) {
    if (!path) throw std::invalid_argument("path is null");
    const std::string name(path);
    Fd f;
    f.fd = ::open(path, O_RDONLY);
    if (f.fd < 0) throw PlyError(name + ": cannot open");
    struct stat sb {};
    if (::fstat(f.fd, &sb) != 0) throw PlyError(name + ": cannot stat");
    const size_t size = static_cast<size_t>(sb.st_size);
    // The header: read until its end_header line is in memory (headers are a few KB).
    std::vector<char> head;
    PlyLayout L;
    for (size_t want = 4096;; want *= 4) {
        head.resize(std::min(want, size));
        size_t got = 0;
        while (got < head.size()) {
            const ssize_t k = ::pread(f.fd, head.data() + got, head.size() - got, static_cast<off_t>(got));
            if (k <= 0) break;
            got += static_cast<size_t>(k);
        }
        head.resize(got);
        const std::string_view text(head.data(), head.size());
        const size_t end = text.find("end_header");
        if ((end != std::string_view::npos && text.find('\n', end) != std::string_view::npos) || got == size ||
            got < std::min(want, size)) {
            L = parsePlyHeader(head.data(), head.size(), name, size);
            break;
        }
    }
    auto r = newRenderer(L);
    const size_t bodyBytes = static_cast<size_t>(L.n) * static_cast<size_t>(L.stride) * 4;
    DevBuf<uint8_t> body;
    body.reserve(std::max<size_t>(bodyBytes, 1));
    if (bodyBytes) {
        hipStream_t st = r->engine->stream;
        std::lock_guard<std::mutex> lock(stagingMutex);
        Staging& stage = staging();
        size_t done = 0;
        try {
            for (int k = 0; done < bodyBytes; k ^= 1) {
                const size_t bytes = std::min(Staging::kBytes, bodyBytes - done);
                hipCheck(hipEventSynchronize(stage.done[k]), "hipEventSynchronize");  // this buffer's last copy
                if (!readAt(f.fd, stage.buf[k], bytes, L.body + done, 4))
                    throw PlyError(name + ": file is shorter than the header promises (" + std::to_string(L.n) +
                                   " Gaussians)");
                hipCheck(hipMemcpyAsync(body.p + done, stage.buf[k], bytes, hipMemcpyHostToDevice, st),
                         "hipMemcpyAsync");
                hipCheck(hipEventRecord(stage.done[k], st), "hipEventRecord");
                done += bytes;
            }
        } catch (...) {
            (void)hipStreamSynchronize(st);  // the staging buffers are not reused while copies read them
            throw;
        }
    }
    return ingestBody(L, reinterpret_cast<const float*>(body.p), std::move(r));
}

ans_renderer* createPly(const void* data, size_t bytes) {
    if (!data) throw std::invalid_argument("data is null");
    const PlyLayout L = parsePlyHeader(static_cast<const char*>(data), bytes, "PLY data", bytes);
    auto r = newRenderer(L);
    const size_t bodyBytes = static_cast<size_t>(L.n) * L.stride * 4;
    DevBuf<uint8_t> body;
    body.reserve(std::max<size_t>(bodyBytes, 1));
    if (bodyBytes)
        hipCheck(hipMemcpyAsync(body.p, static_cast<const uint8_t*>(data) + L.body, bodyBytes, hipMemcpyHostToDevice,
                                r->engine->stream),
                 "hipMemcpyAsync");
    return ingestBody(L, reinterpret_cast<const float*>(body.p), std::move(r));
}

std::string plyHeader(int64_t n, int coeffs) {
    std::string h = "ply\nformat binary_little_endian 1.0\nelement vertex " + std::to_string(n) + "\n";
    auto prop = [&](const std::string& p) { h += "property float " + p + "\n"; };
    for (const char* p : {"x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"}) prop(p);
    for (int k = 0; k < 3 * (coeffs - 1); ++k) prop("f_rest_" + std::to_string(k));
    for (const char* p : {"opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"}) prop(p);
    return h + "end_header\n";
}

}  // namespace

}  // namespace ans

using namespace ans;

extern "C" {

int ans_ply_inspect(const void* data, size_t bytes, int64_t* n, int32_t* sh_degree) {
    return guarded([&] {
        if (!data) throw std::invalid_argument("data is null");
        const PlyLayout L = parsePlyHeader(static_cast<const char*>(data), bytes, "PLY data", bytes);
        if (n) *n = L.n;
        if (sh_degree) *sh_degree = L.degree;
        return ANS_OK;
    });
}

int ans_renderer_create_ply(const void* data, size_t bytes, ans_renderer** out) {
    return guarded([&] {
        if (!out) throw std::invalid_argument("out is null");
        *out = nullptr;
        *out = createPly(data, bytes);
        return ANS_OK;
    });
}

int ans_renderer_load_ply(const char* path, ans_renderer** out) {
    return guarded([&] {
        if (!out) throw std::invalid_argument("out is null");
        *out = nullptr;
        *out = loadPly(path);
        return ANS_OK;
    });
}

int ans_renderer_scene(ans_renderer* r, int64_t* n, int32_t* sh_degree, float* pos_xyz, float* scale_xyz,
                       float* rot_wxyz, float* opacity, float* sh) {
    return guarded([&] {
        if (!r) throw std::invalid_argument("renderer is null");
        const SceneDev& s = r->scene;
        if (n) *n = s.n;
        if (sh_degree) *sh_degree = s.shDegree;
        const size_t m = static_cast<size_t>(s.n);
        auto down = [&](void* dst, const void* src, size_t bytes) {
            if (dst && bytes) hipCheck(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hipMemcpy D2H");
        };
        down(pos_xyz, s.pos, 12 * m);
        down(scale_xyz, s.scale, 12 * m);
        down(rot_wxyz, s.rot, 16 * m);
        down(opacity, s.opacity, 4 * m);
        if (sh && m) {
            const size_t k3 = static_cast<size_t>((s.shDegree + 1) * (s.shDegree + 1)) * 3;
            std::vector<float> planar(k3 * m);
            down(planar.data(), s.sh, 4 * planar.size());
            for (size_t i = 0; i < m; ++i)
                for (size_t j = 0; j < k3; ++j) sh[i * k3 + j] = planar[j * m + i];
        }
        return ANS_OK;
    });
}

int ans_ply_write(const ans_params* params, void* out, size_t capacity, size_t* written, void* stream) {
    return guarded([&] {
        if (!params) throw std::invalid_argument("params is null");
        if (params->n < 0 || params->n > (int64_t(1) << 30))
            throw std::invalid_argument("Gaussian count must be in 0..2^30");
        if (params->sh_degree < 0 || params->sh_degree > 3) throw std::invalid_argument("sh_degree must be in 0..3");
        if (params->activated) throw std::invalid_argument("the PLY layout stores raw parameters (activated must be 0)");
        const int coeffs = (params->sh_degree + 1) * (params->sh_degree + 1);
        const std::string header = plyHeader(params->n, coeffs);
        const int stride = 17 + 3 * (coeffs - 1);
        const size_t bodyBytes = static_cast<size_t>(params->n) * stride * 4;
        const size_t total = header.size() + bodyBytes;
        if (written) *written = total;
        if (!out) return ANS_OK;
        if (capacity < total) throw std::invalid_argument("output buffer is smaller than the PLY file");
        if (params->n > 0 && (!params->means || !params->log_scales || !params->quats || !params->opacity_logits ||
                              !params->sh))
            throw std::invalid_argument("parameter arrays must not be null");
        std::memcpy(out, header.data(), header.size());
        if (bodyBytes) {
            hipStream_t st = static_cast<hipStream_t>(stream);
            DevBuf<float> rows;
            rows.reserve(bodyBytes / 4);
            const uint32_t blocks = static_cast<uint32_t>((params->n + kRows - 1) / kRows);
            k_plyRows<<<blocks, kRows, sizeof(float) * kRows * stride, st>>>(
                params->n, coeffs, params->means, params->log_scales, params->quats, params->opacity_logits, params->sh,
                rows.p);
            hipCheck(hipGetLastError(), "k_plyRows launch");
            hipCheck(hipMemcpyAsync(static_cast<uint8_t*>(out) + header.size(), rows.p, bodyBytes,
                                    hipMemcpyDeviceToHost, st),
                     "hipMemcpyAsync");
            hipCheck(hipStreamSynchronize(st), "hipStreamSynchronize");
        }
        return ANS_OK;
    });
}

}  // extern "C"
