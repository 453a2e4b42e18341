// Files of the reference CLI's verbs around the renderer: COLMAP sparse models in (cameras.bin,
// images.bin; camera.cpp:59-140) and PNG frames out (image.cpp:41-85), the latter also encoded on
// the device for frames that are already in HBM.
#include "aniso_amd.h"
#include "engine.hpp"
#include "png.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <map>
#include <mutex>
#include <string>
#include <vector>

namespace ans {
namespace {

// ---- COLMAP ------------------------------------------------------------------------------

template <typename T>
T readValue(std::istream& in, const std::string& file) {
    T v{};
    in.read(reinterpret_cast<char*>(&v), sizeof(T));
    if (!in) throw std::invalid_argument(file + ": unexpected end of file");
    return v;
}

struct Intrinsic {
    int width = 0, height = 0;
    double fx = 0, fy = 0, cx = 0, cy = 0;
};

// Parameter counts of COLMAP's camera models 0..10 (SIMPLE_PINHOLE, PINHOLE, SIMPLE_RADIAL,
// RADIAL, OPENCV, OPENCV_FISHEYE, FULL_OPENCV, FOV, SIMPLE_RADIAL_FISHEYE, RADIAL_FISHEYE,
// THIN_PRISM_FISHEYE), as COLMAP's binary model format defines them.
int modelParams(int32_t model) {
    static const int n[11] = {3, 4, 4, 5, 8, 8, 12, 5, 4, 5, 12};
    return model >= 0 && model < 11 ? n[model] : -1;
}

// cameras.bin: count (u64), then per camera id (u32), model (i32), width (u64), height (u64),
// params (f64 x modelParams). Only the distortion-free models are accepted. A repeated id keeps
// the last entry. Mirrors camera.cpp:59-90, including its messages.
std::map<uint32_t, Intrinsic> readIntrinsics(const std::string& file) {
    std::ifstream in(file, std::ios::binary);
    if (!in) throw std::invalid_argument(file + ": cannot open");
    std::map<uint32_t, Intrinsic> out;
    const uint64_t count = readValue<uint64_t>(in, file);
    for (uint64_t i = 0; i < count; ++i) {
        const uint32_t id = readValue<uint32_t>(in, file);
        const int32_t model = readValue<int32_t>(in, file);
        Intrinsic c;
        c.width = static_cast<int>(readValue<uint64_t>(in, file));
        c.height = static_cast<int>(readValue<uint64_t>(in, file));
        const int np = modelParams(model);
        if (np < 0) throw std::invalid_argument(file + ": unknown camera model " + std::to_string(model));
        double p[12];
        for (int k = 0; k < np; ++k) p[k] = readValue<double>(in, file);
        if (model == 0) {
            c.fx = c.fy = p[0];
            c.cx = p[1];
            c.cy = p[2];
        } else if (model == 1) {
            c.fx = p[0];
            c.fy = p[1];
            c.cx = p[2];
            c.cy = p[3];
        } else {
            throw std::invalid_argument(file + ": camera model " + std::to_string(model) +
                                        " has lens distortion; undistort the images first");
        }
        out[id] = c;
    }
    return out;
}

// Rotation of the unit quaternion q / |q| (w, x, y, z), in double.
void rotationOf(double w, double x, double y, double z, double* R) {
    const double s = std::sqrt(w * w + x * x + y * y + z * z);
    w /= s, x /= s, y /= s, z /= s;
    R[0] = 1 - 2 * (y * y + z * z), R[1] = 2 * (x * y - w * z), R[2] = 2 * (x * z + w * y);
    R[3] = 2 * (x * y + w * z), R[4] = 1 - 2 * (x * x + z * z), R[5] = 2 * (y * z - w * x);
    R[6] = 2 * (x * z - w * y), R[7] = 2 * (y * z + w * x), R[8] = 1 - 2 * (x * x + y * y);
}

struct NamedCamera {
    std::string name;
    ans_camera cam;
};

// images.bin: count (u64), then per image id (u32), qvec (f64 x 4, w x y z), tvec (f64 x 3),
// camera id (u32), NUL-terminated name, 2D point count (u64) and 24 bytes per point (skipped).
// Cameras sorted by image name. Mirrors camera.cpp:104-140.
std::vector<NamedCamera> readColmap(const std::string& dir) {
    const std::string sep = (!dir.empty() && dir.back() == '/') ? "" : "/";
    const auto intr = readIntrinsics(dir + sep + "cameras.bin");
    const std::string file = dir + sep + "images.bin";
    std::ifstream in(file, std::ios::binary);
    if (!in) throw std::invalid_argument(file + ": cannot open");
    std::vector<NamedCamera> out;
    const uint64_t count = readValue<uint64_t>(in, file);
    for (uint64_t i = 0; i < count; ++i) {
        (void)readValue<uint32_t>(in, file);
        double q[4], t[3];
        for (double& v : q) v = readValue<double>(in, file);
        for (double& v : t) v = readValue<double>(in, file);
        const uint32_t camId = readValue<uint32_t>(in, file);
        std::string name;
        for (char ch = readValue<char>(in, file); ch != '\0'; ch = readValue<char>(in, file)) name.push_back(ch);
        const uint64_t points = readValue<uint64_t>(in, file);
        in.seekg(static_cast<std::streamoff>(points * 24), std::ios::cur);
        const auto it = intr.find(camId);
        if (it == intr.end()) throw std::invalid_argument(file + ": image " + name + " refers to a missing camera");
        NamedCamera nc;
        nc.name = std::move(name);
        ans_camera& c = nc.cam;
        c = ans_camera{};
        rotationOf(q[0], q[1], q[2], q[3], c.R);
        for (int k = 0; k < 3; ++k) c.t[k] = t[k];
        c.fx = it->second.fx, c.fy = it->second.fy, c.cx = it->second.cx, c.cy = it->second.cy;
        c.width = it->second.width, c.height = it->second.height;
        out.push_back(std::move(nc));
    }
    std::stable_sort(out.begin(), out.end(), [](const NamedCamera& a, const NamedCamera& b) { return a.name < b.name; });
    return out;
}

// ---- PNG on the device ---------------------------------------------------------------------

constexpr int kMaxDevices = 64;
std::mutex g_zMutex;
bool g_zReady[kMaxDevices] = {};

// Uploads the CRC shift matrices to the current device's constant memory once.
void zShiftReady() {
    int device = 0;
    hipCheck(hipGetDevice(&device), "hipGetDevice");
    if (device < 0 || device >= kMaxDevices) throw std::runtime_error("device index out of range");
    std::lock_guard<std::mutex> lock(g_zMutex);
    if (g_zReady[device]) return;
    hipCheck(hipMemcpyToSymbol(HIP_SYMBOL(png::cZ), png::shiftTable().m, sizeof(png::ShiftTable), 0,
                               hipMemcpyHostToDevice),
             "hipMemcpyToSymbol");
    g_zReady[device] = true;
}

void checkSize(int w, int h) {
    if (w < 0 || h < 0) throw std::invalid_argument("negative image size");
    // The IDAT chunk's length field is 32 bits (and the CRC shifts cover lengths below 2^40).
    const uint64_t raw = static_cast<uint64_t>(h) * (3 * static_cast<uint64_t>(w) + 1);
    if (raw + 5 * (raw / png::kStoredMax + 1) + 6 > 0xffffffffull)
        throw std::invalid_argument("image too large for one PNG IDAT chunk");
}

}  // namespace

// Enqueues the device encoder of a w x h RGB8 frame on st: out receives png::Layout(w, h).total
// bytes; work holds png::DevWork(layout).bytes.
void enqueuePng(const uint8_t* rgb, int w, int h, uint8_t* out, uint8_t* work, hipStream_t st) {
    const png::Layout L(w, h);
    const png::DevWork W(L);
    const png::Frame f = png::frameBytes(w, h, L);
    const png::DevLayout D{L.stride, L.raw, L.blocks, L.zlib, L.idatData};
    uint32_t* part = reinterpret_cast<uint32_t*>(work + W.adlerOff);
    uint32_t* reg = reinterpret_cast<uint32_t*>(work + W.regOff);
    uint64_t* len = reinterpret_cast<uint64_t*>(work + W.lenOff);
    if (L.raw > 0) png::k_pngBody<<<static_cast<uint32_t>(W.bodyBlocks), png::kBodyBlock, 0, st>>>(rgb, out, D, part);
    png::k_pngAdler<<<1, 256, 0, st>>>(out, D, f, part, L.raw > 0 ? static_cast<uint32_t>(W.bodyBlocks) : 0u);
    png::k_crcPieces<<<static_cast<uint32_t>(W.crcBlocks), 256, 0, st>>>(out + L.idatData - 4, L.zlib + 4, reg, len);
    png::k_crcFinish<<<1, 256, 0, st>>>(reg, len, static_cast<uint32_t>(W.crcBlocks), out + L.idatData + L.zlib);
    hipCheck(hipGetLastError(), "png kernel launch");
}

}  // namespace ans

using namespace ans;

extern "C" {

int ans_colmap_load(const char* sparse_dir, ans_camera* cameras, char* names, size_t name_capacity, int capacity,
                    int* count) {
    return guarded([&] {
        if (!sparse_dir) throw std::invalid_argument("sparse_dir is null");
        if (!count) throw std::invalid_argument("count is null");
        const auto cams = readColmap(sparse_dir);
        *count = static_cast<int>(cams.size());
        if (!cameras) return ANS_OK;
        if (capacity < static_cast<int>(cams.size())) {
            g_lastError = "the model holds " + std::to_string(cams.size()) + " images, capacity is " +
                          std::to_string(capacity);
            return ANS_ERR_CAPACITY;
        }
        for (size_t i = 0; i < cams.size(); ++i) {
            cameras[i] = cams[i].cam;
            if (names) {
                if (cams[i].name.size() + 1 > name_capacity)
                    throw std::invalid_argument("image name '" + cams[i].name + "' does not fit name_capacity");
                std::memcpy(names + i * name_capacity, cams[i].name.c_str(), cams[i].name.size() + 1);
            }
        }
        return ANS_OK;
    });
}

int ans_png_encode(const uint8_t* rgb, int width, int height, void* out, size_t capacity, size_t* written) {
    return guarded([&] {
        checkSize(width, height);
        if (!written) throw std::invalid_argument("written is null");
        const png::Layout L(width, height);
        *written = L.total;
        if (!out) return ANS_OK;
        if (capacity < L.total) throw std::invalid_argument("capacity below the PNG's size");
        if (!rgb && L.raw > static_cast<size_t>(height)) throw std::invalid_argument("rgb is null");
        png::encodeHost(rgb, width, height, static_cast<uint8_t*>(out));
        return ANS_OK;
    });
}

int ans_png_write(const char* path, const uint8_t* rgb, int width, int height) {
    return guarded([&] {
        if (!path) throw std::invalid_argument("path is null");
        checkSize(width, height);
        if (!rgb && width > 0 && height > 0) throw std::invalid_argument("rgb is null");
        const png::Layout L(width, height);
        std::vector<uint8_t> bytes(L.total);
        png::encodeHost(rgb, width, height, bytes.data());
        FILE* f = std::fopen(path, "wb");
        if (!f) throw std::invalid_argument(std::string(path) + ": cannot write");
        const size_t n = std::fwrite(bytes.data(), 1, bytes.size(), f);
        const int closed = std::fclose(f);
        if (n != bytes.size() || closed != 0) throw std::invalid_argument(std::string(path) + ": cannot write");
        return ANS_OK;
    });
}

int ans_png_device_sizes(int width, int height, size_t* file_bytes, size_t* workspace_bytes) {
    return guarded([&] {
        checkSize(width, height);
        const png::Layout L(width, height);
        if (file_bytes) *file_bytes = L.total;
        if (workspace_bytes) *workspace_bytes = png::DevWork(L).bytes;
        return ANS_OK;
    });
}

int ans_png_encode_device(const void* rgb, int width, int height, void* out, size_t capacity, void* workspace,
                          size_t workspace_bytes, void* stream) {
    return guarded([&] {
        checkSize(width, height);
        if (width == 0 || height == 0) throw std::invalid_argument("the device encoder takes non-empty images");
        if (!rgb || !out || !workspace) throw std::invalid_argument("null device pointer");
        const png::Layout L(width, height);
        if (capacity < L.total) throw std::invalid_argument("capacity below the PNG's size");
        if (workspace_bytes < png::DevWork(L).bytes) throw std::invalid_argument("workspace too small");
        zShiftReady();
        enqueuePng(static_cast<const uint8_t*>(rgb), width, height, static_cast<uint8_t*>(out),
                   static_cast<uint8_t*>(workspace), static_cast<hipStream_t>(stream));
        return ANS_OK;
    });
}

}  // extern "C"
