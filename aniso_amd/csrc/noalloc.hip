// The differentiable forward's image without any allocation: every per-frame buffer of the
// launch sequence (splat records, sort and scan scratch, pair buffers, tile ranges, control
// block) is laid out in one fixed array in device memory, `scratch`, that the library holds for
// the life of the process on each device it runs on. For callers that cannot allocate (code
// under HIP graph capture, tight serving loops). No counterpart in the reference.
#include "aniso_amd.h"
#include "engine.hpp"

#include <mutex>
#include <stdexcept>
#include <string>

namespace ans {

namespace {

constexpr size_t kScratchBytes = size_t(64) << 20;  // 64 MiB per device
constexpr int64_t kMaxGaussians = int64_t(1) << 19;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

// The device address of `scratch` on each device, looked up once per device with that device
// current (each device holds its own copy of the array).
std::mutex g_scratchMutex;
uint8_t* g_scratchPtr[kMaxDevices] = {};

uint8_t* scratchOf(int device) {
    if (device < 0 || device >= kMaxDevices) throw std::runtime_error("device index out of range");
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    if (!g_scratchPtr[device]) {
        void* ptr = nullptr;
        hipCheck(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress");
        if (!ptr) throw std::runtime_error("no device address for the scratch array");
        g_scratchPtr[device] = static_cast<uint8_t*>(ptr);
    }
    return g_scratchPtr[device];
}

// Carves FrameBuffers for n Gaussians and `tiles` tiles out of the scratch array; the pair
// buffers take what is left. Returns false if the per-Gaussian part does not fit.
bool layout(uint8_t* base, size_t n, size_t tiles, FrameBuffers& b) {
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t at = off;
        off = (off + bytes + 255) / 256 * 256;
        return base + at;
    };
    const size_t m = std::max<size_t>(n, 1);
    b.ctl = reinterpret_cast<FrameCtl*>(take(sizeof(FrameCtl)));
    b.rec = reinterpret_cast<SplatRec*>(take(sizeof(SplatRec) * m));
    b.tilesTouched = reinterpret_cast<uint32_t*>(take(4 * m));
    b.order = reinterpret_cast<uint32_t*>(take(4 * m));
    b.depthKey = reinterpret_cast<uint32_t*>(take(4 * m));
    b.sortKA = reinterpret_cast<uint32_t*>(take(4 * m));
    b.sortVA = reinterpret_cast<uint32_t*>(take(4 * m));
    b.sortKB = reinterpret_cast<uint32_t*>(take(4 * m));
    b.sortVB = reinterpret_cast<uint32_t*>(take(4 * m));
    b.ttByRank = reinterpret_cast<uint32_t*>(take(4 * m));
    b.rectIdx = reinterpret_cast<uint2*>(take(8 * m));
    b.rectByRank = reinterpret_cast<uint2*>(take(8 * m));
    const size_t blocksDup = (m + kDupTile - 1) / kDupTile;
    b.blockPairs = reinterpret_cast<uint32_t*>(take(4 * blocksDup));
    b.blockVisible = reinterpret_cast<uint32_t*>(take(4 * blocksDup));
    b.rowTotal = reinterpret_cast<uint32_t*>(take(4 * kRadix));
    b.ranges = reinterpret_cast<uint2*>(take(8 * std::max<size_t>(tiles, 1)));
    b.pixState = nullptr;
    if (off >= kScratchBytes) return false;
    // The rest: histograms (kRadix words per sort block of either sort) and four pair arrays.
    const size_t blocksDepth = (m + kSortBlock * kDepthItems - 1) / (kSortBlock * kDepthItems);
    const size_t perPair = 16;                                      // tile and Gaussian, double-buffered
    const size_t histPerPair = kRadix * 4 / (kSortBlock * kTileItems);  // a fraction of a byte, rounded below
    const size_t left = kScratchBytes - off;
    size_t cap = (left - std::min(left, size_t(4) * kRadix * (blocksDepth + 1) + 5 * 256)) / (perPair + histPerPair + 1);
    cap = cap / 256 * 256;
    if (cap == 0) return false;
    const size_t blocksPair = (cap + kSortBlock * kTileItems - 1) / (kSortBlock * kTileItems);
    b.hist = reinterpret_cast<uint32_t*>(take(4 * kRadix * std::max({blocksDepth, blocksPair, size_t(1)})));
    b.tileA = reinterpret_cast<uint32_t*>(take(4 * cap));
    b.gaussA = reinterpret_cast<uint32_t*>(take(4 * cap));
    b.tileB = reinterpret_cast<uint32_t*>(take(4 * cap));
    b.gaussB = reinterpret_cast<uint32_t*>(take(4 * cap));
    b.pairCap = cap;
    return off <= kScratchBytes;
}

}  // namespace

SceneDev sceneOfParams(const ans_params* p);

}  // namespace ans

using namespace ans;

extern "C" {

int ans_diff_forward_noalloc(const ans_params* params, const ans_camera* cam, const ans_options* options,
                             float* image, void* stream, ans_stats* stats) {
    return guarded([&] {
        checkCamera(cam);
        const SceneDev s = sceneOfParams(params);
        if (s.n > kMaxGaussians)
            throw std::invalid_argument("the no-allocation forward takes at most 524288 Gaussians");
        if (!image) throw std::invalid_argument("image is null");
        const ans_options o = resolveOptions(options);
        const CamParams cp = makeCamParams(*cam, o, s.shDegree);
        if (cp.tilesX > 65535 || cp.tilesY > 65535) throw std::invalid_argument("too many tiles");
        const size_t tiles = static_cast<size_t>(cp.tilesX) * static_cast<size_t>(cp.tilesY);
        int device = 0;
        hipCheck(hipGetDevice(&device), "hipGetDevice");
        FrameBuffers b{};
        if (!layout(scratchOf(device), static_cast<size_t>(s.n), tiles, b))
            throw std::invalid_argument("scene and image do not fit the no-allocation scratch (64 MiB)");
        hipStream_t st = static_cast<hipStream_t>(stream);
        enqueueCountKernels(b, s, cp, st, nullptr, nullptr, 0);
        const uint32_t* sortedTile = nullptr;
        const uint32_t* sortedGauss = nullptr;
        enqueueRestKernels(b, static_cast<uint32_t>(s.n), cp, o, false, nullptr, image, st, nullptr, nullptr,
                           &sortedTile, &sortedGauss);
        // The pair count decides whether the frame fit: the kernels after the count skip their
        // work if it did not.
        FrameCtl ctl{};
        hipCheck(hipMemcpyAsync(&ctl, b.ctl, sizeof(FrameCtl), hipMemcpyDeviceToHost, st), "hipMemcpyAsync ctl");
        hipCheck(hipStreamSynchronize(st), "hipStreamSynchronize");
        const uint32_t pairs = s.n ? ctl.pairs : 0, visible = s.n ? ctl.visible : 0;
        if (stats) {
            *stats = ans_stats{};
            stats->visible = visible;
            stats->tile_pairs = pairs;
            stats->pair_capacity = static_cast<int64_t>(b.pairCap);
        }
        if (s.n && ctl.overflow) {
            g_lastError = "pair count " + std::to_string(pairs) + " exceeds the no-allocation scratch's " +
                          std::to_string(b.pairCap) + " pairs";
            return ANS_ERR_CAPACITY;
        }
        return ANS_OK;
    });
}

}  // extern "C"
