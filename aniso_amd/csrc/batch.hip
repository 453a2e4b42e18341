// Batched multi-view forward: many cameras of one scene rendered by one sequence of batched
// launches (blockIdx.y = view, or (view, tile) work items for the blend), replayed as a captured
// HIP graph. There is no batched renderer in the reference; its nearest caller renders a camera
// path one frame at a time (src/main.cpp:234-244), each frame a full GpuRenderer::render with its
// own host round trip (src/gpu_render.cu:324-329).
//
// Per view the pipeline is the single-frame one (views.hpp, forward.hpp, sort.hpp) run on that
// view's segment of a shared arena, so each view's splat records, lists, ranges and image equal
// a single-frame render of that camera with the same options.
#include "aniso_amd.h"
#include "batch.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

namespace ans {

ViewBatch::ViewBatch() {
    hipCheck(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
    for (auto& slot : ev)
        for (auto& e : slot) hipCheck(hipEventCreate(&e), "hipEventCreate");
    for (int k = 0; k < 2; ++k) {
        hipCheck(hipHostMalloc(reinterpret_cast<void**>(&tableHost[k]), TableLayout(kMaxViews).bytes,
                               hipHostMallocDefault),
                 "hipHostMalloc");
        hipCheck(hipHostMalloc(reinterpret_cast<void**>(&ctlHost[k]), sizeof(FrameCtl) * (kMaxViews + 1),
                               hipHostMallocDefault),
                 "hipHostMalloc");
    }
    hipCheck(hipHostMalloc(reinterpret_cast<void**>(&mirror), sizeof(CountMirror) * kMaxViews,
                           hipHostMallocMapped | hipHostMallocCoherent),
             "hipHostMalloc");
    std::memset(mirror, 0, sizeof(CountMirror) * kMaxViews);
    hipCheck(hipHostGetDevicePointer(reinterpret_cast<void**>(&mirrorDev), mirror, 0), "hipHostGetDevicePointer");
}

ViewBatch::~ViewBatch() {
    if (pathRender) (void)hipStreamSynchronize(pathRender);
    if (pathCopy) (void)hipStreamSynchronize(pathCopy);
    // Kernels of a submitted batch may still write the pinned count mirror.
    for (auto& slot : ev)
        if (slot[2]) (void)hipEventSynchronize(slot[2]);
    dropGraphs();
    if (mirror) (void)hipHostFree(mirror);
    for (int k = 0; k < 2; ++k) {
        if (tableHost[k]) (void)hipHostFree(tableHost[k]);
        if (ctlHost[k]) (void)hipHostFree(ctlHost[k]);
        if (pathHost[k]) (void)hipHostFree(pathHost[k]);
        if (copied[k]) (void)hipEventDestroy(copied[k]);
    }
    for (auto& slot : ev)
        for (auto& e : slot)
            if (e) (void)hipEventDestroy(e);
    if (pathRender) (void)hipStreamDestroy(pathRender);
    if (pathCopy) (void)hipStreamDestroy(pathCopy);
    if (stream) (void)hipStreamDestroy(stream);
}

void ViewBatch::dropGraphs() {
    for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
    graphs.clear();
}

void ViewBatchDeleter::operator()(ViewBatch* b) const { delete b; }

namespace {

// Whether buffer b must grow to n elements; with APPLY it grows.
template <bool APPLY, class T>
bool grow(DevBuf<T>& b, size_t n) {
    if (n <= b.cap) return false;
    if (APPLY) b.reserve(n);
    return true;
}

// Whether any arena buffer must grow for the batch L; with APPLY they grow.
template <bool APPLY>
bool arenaGrowth(ViewBatch& B, const Layout& L) {
    const size_t V = L.views, n = std::max<uint32_t>(L.n, 1), vn = V * n;
    bool moved = false;
    moved |= grow<APPLY>(B.table, TableLayout(L.views).bytes);
    moved |= grow<APPLY>(B.ctl, V + 1);
    moved |= grow<APPLY>(B.rec, vn);
    moved |= grow<APPLY>(B.depthKey, vn);
    moved |= grow<APPLY>(B.order, vn);
    moved |= grow<APPLY>(B.kA, vn);
    moved |= grow<APPLY>(B.vA, vn);
    moved |= grow<APPLY>(B.kB, vn);
    moved |= grow<APPLY>(B.vB, vn);
    moved |= grow<APPLY>(B.ttByRank, vn);
    moved |= grow<APPLY>(B.rectIdx, vn);
    moved |= grow<APPLY>(B.rectByRank, vn);
    const size_t histD = V * kRadix * size_t(L.blocksD), histP = V * kRadix * size_t(L.blocksP),
                 histI = size_t(kRadix) * L.blocksI;
    moved |= grow<APPLY>(B.hist, std::max({histD, histP, histI}));
    moved |= grow<APPLY>(B.rowTotal, (V + 1) * kRadix);
    moved |= grow<APPLY>(B.blockPairs, V * L.blocksDup);
    moved |= grow<APPLY>(B.blockVisible, V * L.blocksDup);
    const size_t vp = V * L.pairCap;
    moved |= grow<APPLY>(B.tileA, vp);
    moved |= grow<APPLY>(B.gaussA, vp);
    moved |= grow<APPLY>(B.tileB, vp);
    moved |= grow<APPLY>(B.gaussB, vp);
    moved |= grow<APPLY>(B.ranges, size_t(L.items));
    moved |= grow<APPLY>(B.itemKey, L.items);
    moved |= grow<APPLY>(B.itemK2, L.items);
    moved |= grow<APPLY>(B.itemV2, L.items);
    moved |= grow<APPLY>(B.itemOrder, L.items);
    if (L.diff) {
        moved |= grow<APPLY>(B.tilesTouched, vn);
        moved |= grow<APPLY>(B.pixState, std::max<uint64_t>(L.pixels, 1));
    }
    return moved;
}

}  // namespace

bool arenaFits(ViewBatch& B, const Layout& L) { return !arenaGrowth<false>(B, L); }

void reserveArena(ViewBatch& B, const Layout& L) {
    if (arenaGrowth<true>(B, L)) {
        B.dropGraphs();
        ++B.generation;
        // A fresh control array: zero it once (the per-call init leaves sticky alone).
        hipCheck(hipMemset(B.ctl.p, 0, sizeof(FrameCtl) * (L.views + 1)), "hipMemset ctl");
    }
}

namespace {

template <int TW, int TH, int BATCH>
void launchBlendViews(ViewBatch& B, const Layout& L, const TableLayout& T, const uint32_t* sortedGauss,
                      hipStream_t st) {
    const CamParams* cams = reinterpret_cast<const CamParams*>(B.table.p + T.cams);
    uint8_t* const* outs = reinterpret_cast<uint8_t* const*>(B.table.p + T.outs);
    float* const* outf = reinterpret_cast<float* const*>(B.table.p + T.outf);
    const uint64_t* pixOff = reinterpret_cast<const uint64_t*>(B.table.p + T.pixOff);
    const uint32_t pairStride = static_cast<uint32_t>(L.pairCap);
    if (L.diff)  // frames for the backward pass always cull (culling changes no pixel)
        k_blendViews<TW, TH, BATCH, true, true><<<L.totalTiles, TW * TH, 0, st>>>(
            B.itemOrder.p, L.maxTiles, B.ranges.p, sortedGauss, pairStride, B.rec.p, L.n, cams, outs, outf,
            B.pixState.p, pixOff, B.ctl.p);
    else if (L.cull)
        k_blendViews<TW, TH, BATCH, true, false><<<L.totalTiles, TW * TH, 0, st>>>(
            B.itemOrder.p, L.maxTiles, B.ranges.p, sortedGauss, pairStride, B.rec.p, L.n, cams, outs, outf, nullptr,
            pixOff, B.ctl.p);
    else
        k_blendViews<TW, TH, BATCH, false, false><<<L.totalTiles, TW * TH, 0, st>>>(
            B.itemOrder.p, L.maxTiles, B.ranges.p, sortedGauss, pairStride, B.rec.p, L.n, cams, outs, outf, nullptr,
            pixOff, B.ctl.p);
}

template <int TW, int TH>
void launchBlendViewsT(ViewBatch& B, const Layout& L, const TableLayout& T, const uint32_t* sortedGauss,
                       hipStream_t st) {
    if (L.opts.batch_depth == 64) launchBlendViews<TW, TH, 64>(B, L, T, sortedGauss, st);
    else if (L.opts.batch_depth == 128) launchBlendViews<TW, TH, 128>(B, L, T, sortedGauss, st);
    else launchBlendViews<TW, TH, 256>(B, L, T, sortedGauss, st);
}

}  // namespace

const uint32_t* enqueueBatch(ViewBatch& B, const SceneDev& s, const Layout& L, hipStream_t st) {
    const TableLayout T(L.views);
    const CamParams* cams = reinterpret_cast<const CamParams*>(B.table.p + T.cams);
    const int* tilesX = reinterpret_cast<const int*>(B.table.p + T.tilesX);
    const int* tilesN = reinterpret_cast<const int*>(B.table.p + T.tiles);
    const uint32_t V = static_cast<uint32_t>(L.views), n = L.n;
    const uint32_t cap = static_cast<uint32_t>(L.pairCap);
    FrameCtl* ctl = B.ctl.p;

    k_initViews<<<(std::max(V + 1, L.items) + 255) / 256, 256, 0, st>>>(ctl, V + 1, B.ranges.p, L.items);
    if (n > 0) {
        if (L.diff)
            k_preprocessViewsRaw<<<(n + 255) / 256, 256, 0, st>>>(
                static_cast<int>(n), L.views, reinterpret_cast<const DiffParamsDev*>(B.table.p + T.params), cams,
                B.rec.p, B.tilesTouched.p, B.depthKey.p, B.rectIdx.p);
        else
            k_preprocessViews<<<(n + 255) / 256, 256, 0, st>>>(static_cast<int>(n), L.views, s.pos, s.scale, s.rot,
                                                               s.opacity, s.sh, (s.shDegree + 1) * (s.shDegree + 1),
                                                               cams, B.rec.p, nullptr, B.depthKey.p, B.rectIdx.p);
        // Depth sort of each view's Gaussians: four stable 8-bit passes.
        const uint32_t histD = kRadix * L.blocksD;  // depth sort: kPairItems keys per thread
        const uint32_t* kin = B.depthKey.p;
        const uint32_t* vin = nullptr;
        uint32_t* kout[4] = {B.kA.p, B.kB.p, B.kA.p, nullptr};
        uint32_t* vout[4] = {B.vA.p, B.vB.p, B.vA.p, B.order.p};
        const dim3 gD(L.blocksD, V);
        // The first pass drops the culled Gaussians (key 0xffffffff: a quarter of the
        // Gaussian-views of a typical batch), so the later passes sort only the visible ones.
        for (int pass = 0; pass < 4; ++pass) {
            const int from = pass == 0 ? kCountDrop : kCountKept;
            k_radixUpsweep<kPairItems><<<gD, kSortBlock, 0, st>>>(kin, from, n, n, 8 * pass, 8, B.hist.p, L.blocksD,
                                                                  ctl, n, histD);
            k_scanRows<<<dim3(kRadix, V), kSortBlock, 0, st>>>(B.hist.p, L.blocksD, L.blocksD, B.rowTotal.p, ctl,
                                                               from, n, kSortBlock * kPairItems, histD);
            k_radixDownsweep<kPairItems><<<gD, kSortBlock, 0, st>>>(kin, vin, kout[pass], vout[pass], from, n, n,
                                                                     8 * pass, 8, B.hist.p, B.rowTotal.p, L.blocksD,
                                                                     ctl, n, histD);
            kin = kout[pass];
            vin = vout[pass];
        }
        // Pairs of each view in depth order, into the view's region of the pair arena.
        const dim3 gDup(L.blocksDup, V);
        k_pairTotals<<<gDup, kDupBlock, 0, st>>>(B.order.p, B.rectIdx.p, n, B.ttByRank.p, B.rectByRank.p, B.blockPairs.p,
                                                 B.blockVisible.p, n, L.blocksDup, nullptr, ctl);
        k_scanPairs<<<dim3(1, V), kScanBlock, 0, st>>>(B.blockPairs.p, B.blockVisible.p, L.blocksDup, cap, ctl,
                                                      L.blocksDup, L.diff ? B.mirrorDev : nullptr, 1u);
        k_expandPairs<<<gDup, kDupBlock, 0, st>>>(B.order.p, B.ttByRank.p, B.rectByRank.p, n, 0, B.blockPairs.p,
                                                  B.tileA.p, B.gaussA.p, ctl, n, L.blocksDup, cap, tilesX);
    }
    // Stable sort of each view's pairs by tile id.
    uint32_t* tk[2] = {B.tileA.p, B.tileB.p};
    uint32_t* tv[2] = {B.gaussA.p, B.gaussB.p};
    const uint32_t histP = kRadix * L.blocksP;
    const dim3 gP(L.blocksP, V);
    if (n > 0) {
        for (int pass = 0; pass < L.passes; ++pass) {
            const int src = pass & 1, dst = src ^ 1;
            const int shift = pass * L.digitBits;
            const int bits = std::min(L.digitBits, L.tileBits - shift);
            k_radixUpsweep<kPairItems><<<gP, kSortBlock, 0, st>>>(tk[src], true, 0, cap, shift, bits, B.hist.p,
                                                                  L.blocksP, ctl, cap, histP);
            k_scanRows<<<dim3(1u << bits, V), kSortBlock, 0, st>>>(B.hist.p, L.blocksP, L.blocksP, B.rowTotal.p, ctl,
                                                                   true, cap, kSortBlock * kPairItems, histP);
            k_radixDownsweep<kPairItems><<<gP, kSortBlock, 0, st>>>(tk[src], tv[src], tk[dst], tv[dst], true, 0, cap,
                                                                    shift, bits, B.hist.p, B.rowTotal.p, L.blocksP,
                                                                    ctl, cap, histP);
        }
    }
    const uint32_t* sortedTile = tk[L.passes & 1];
    const uint32_t* sortedGauss = tv[L.passes & 1];
    if (n > 0) {
        const uint32_t rangeBlocks = std::max<uint32_t>(1, std::min<uint32_t>((cap + 1023) / 1024, 2048));
        k_tileRanges<<<dim3(rangeBlocks, V), 256, 0, st>>>(sortedTile, ctl, B.ranges.p, cap, L.maxTiles);
    }
    // Work items longest list first: a stable sort of (view, tile) items by a 16-bit key.
    k_itemKeys<<<(L.items + 255) / 256, 256, 0, st>>>(B.ranges.p, tilesN, L.maxTiles, L.items, B.itemKey.p);
    {
        FrameCtl* ictl = ctl + V;  // never overflows
        const uint32_t* kin = B.itemKey.p;
        const uint32_t* vin = nullptr;
        uint32_t* kout[2] = {B.itemK2.p, nullptr};
        uint32_t* vout[2] = {B.itemV2.p, B.itemOrder.p};
        for (int pass = 0; pass < 2; ++pass) {
            k_radixUpsweep<kDepthItems><<<L.blocksI, kSortBlock, 0, st>>>(kin, false, L.items, L.items, 8 * pass, 8,
                                                                          B.hist.p, L.blocksI, ictl, 0, 0);
            k_scanRows<<<kRadix, kSortBlock, 0, st>>>(B.hist.p, L.blocksI, L.blocksI, B.rowTotal.p, ictl, false,
                                                      L.items, kSortBlock * kDepthItems, 0);
            k_radixDownsweep<kDepthItems><<<L.blocksI, kSortBlock, 0, st>>>(kin, vin, kout[pass], vout[pass], false,
                                                                            L.items, L.items, 8 * pass, 8, B.hist.p,
                                                                            B.rowTotal.p, L.blocksI, ictl, 0, 0);
            kin = kout[pass];
            vin = vout[pass];
        }
    }
    if (L.opts.tile_w == 8) launchBlendViewsT<8, 8>(B, L, T, sortedGauss, st);
    else if (L.opts.tile_w == 32) launchBlendViewsT<32, 16>(B, L, T, sortedGauss, st);
    else launchBlendViewsT<16, 16>(B, L, T, sortedGauss, st);
    hipCheck(hipGetLastError(), "kernel launch");
    return sortedGauss;
}

uint32_t tileKey(const ans_options& o) {
    return static_cast<uint32_t>(o.tile_w) << 16 | static_cast<uint32_t>(o.tile_h);
}

Layout makeLayout(const ViewBatch& B, const SceneDev& s, int views, const std::vector<CamParams>& cps,
                  const ans_options& o, bool diff) {
    Layout L{};
    L.views = views;
    L.n = static_cast<uint32_t>(s.n);
    L.opts = o;
    L.cull = !(o.flags & ANS_FLAG_NO_CULL);
    L.diff = diff;
    uint64_t total = 0;
    uint32_t maxT = 1;
    for (const CamParams& cp : cps) {
        const uint32_t t = static_cast<uint32_t>(cp.tilesX) * static_cast<uint32_t>(cp.tilesY);
        maxT = std::max(maxT, t);
        total += t;
        L.pixels += static_cast<uint64_t>(cp.width) * static_cast<uint64_t>(cp.height);
    }
    if (uint64_t(maxT) * views >= (uint64_t(1) << 31)) throw std::invalid_argument("too many tiles in the batch");
    L.maxTiles = maxT;
    L.totalTiles = static_cast<uint32_t>(total);
    L.items = maxT * static_cast<uint32_t>(views);
    L.tileBits = std::max(1, bitsFor(maxT - 1));
    L.passes = (L.tileBits + 7) / 8;
    L.digitBits = (L.tileBits + L.passes - 1) / L.passes;
    L.blocksD = std::max<uint32_t>(1, (L.n + kSortBlock * kPairItems - 1) / (kSortBlock * kPairItems));
    L.blocksDup = std::max<uint32_t>(1, (L.n + kDupTile - 1) / kDupTile);
    const auto cap = B.capByTile.find(tileKey(o));
    L.pairCap = cap != B.capByTile.end() ? cap->second : std::max<size_t>(size_t(L.n) * 8, size_t(1) << 20);
    L.blocksP = static_cast<uint32_t>((L.pairCap + kSortBlock * kPairItems - 1) / (kSortBlock * kPairItems));
    L.blocksI = std::max<uint32_t>(1, (L.items + kSortBlock * kDepthItems - 1) / (kSortBlock * kDepthItems));
    return L;
}

// Graphs are captured on the batch's own stream and launched on st.
void submitBatch(ViewBatch& B, const SceneDev& s, const Layout& L, const std::vector<CamParams>& cps,
                 const BatchIO& io, hipStream_t st, int slot) {
    const TableLayout T(L.views);
    uint8_t* th = B.tableHost[slot];
    uint64_t pix = 0;
    for (int v = 0; v < L.views; ++v) {
        reinterpret_cast<CamParams*>(th + T.cams)[v] = cps[v];
        reinterpret_cast<uint8_t**>(th + T.outs)[v] = io.outs ? io.outs[v] : nullptr;
        reinterpret_cast<float**>(th + T.outf)[v] = io.outf ? io.outf[v] : nullptr;
        reinterpret_cast<uint64_t*>(th + T.pixOff)[v] = pix;
        pix += static_cast<uint64_t>(cps[v].width) * static_cast<uint64_t>(cps[v].height);
        reinterpret_cast<int*>(th + T.tilesX)[v] = cps[v].tilesX;
        reinterpret_cast<int*>(th + T.tiles)[v] = cps[v].tilesX * cps[v].tilesY;
        reinterpret_cast<const float**>(th + T.grads)[v] = nullptr;
        reinterpret_cast<uint64_t*>(th + T.pairBase)[v] = 0;
    }
    *reinterpret_cast<DiffParamsDev*>(th + T.params) = io.params ? *io.params : DiffParamsDev{};
    if (L.diff)  // the scan sets seq to 1 when it publishes the view's count
        for (int v = 0; v < L.views; ++v) __atomic_store_n(&B.mirror[v].seq, 0u, __ATOMIC_RELEASE);
    hipCheck(hipMemcpyAsync(B.table.p, th, T.bytes, hipMemcpyHostToDevice, st), "hipMemcpyAsync table");
    hipCheck(hipEventRecord(B.ev[slot][0], st), "hipEventRecord");
    if (L.opts.flags & ANS_FLAG_NO_GRAPH) {
        B.sortedSlot[slot] = enqueueBatch(B, s, L, st);
    } else {
        const ViewBatch::GraphKey key{L.views,    s.n,          L.opts.tile_w, L.opts.tile_h,
                                      L.opts.batch_depth, L.cull, L.maxTiles, L.totalTiles,
                                      L.pairCap,  B.generation, L.diff,      L.pixels};
        auto it = B.graphs.find(key);
        if (it == B.graphs.end()) {
            hipGraph_t g = nullptr;
            hipCheck(hipStreamBeginCapture(B.stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
            try {
                enqueueBatch(B, s, L, B.stream);
            } catch (...) {
                (void)hipStreamEndCapture(B.stream, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            hipCheck(hipStreamEndCapture(B.stream, &g), "hipStreamEndCapture");
            hipGraphExec_t exec = nullptr;
            const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            hipCheck(e, "hipGraphInstantiate");
            it = B.graphs.emplace(key, exec).first;
        }
        hipCheck(hipGraphLaunch(it->second, st), "hipGraphLaunch");
        B.sortedSlot[slot] = (L.passes & 1) ? B.gaussB.p : B.gaussA.p;
    }
    hipCheck(hipEventRecord(B.ev[slot][1], st), "hipEventRecord");
    hipCheck(hipMemcpyAsync(B.ctlHost[slot], B.ctl.p, sizeof(FrameCtl) * L.views, hipMemcpyDeviceToHost, st),
             "hipMemcpyAsync ctl");
    hipCheck(hipEventRecord(B.ev[slot][2], st), "hipEventRecord");
}

bool finishBatch(ViewBatch& B, const Layout& L, const std::vector<CamParams>& cps, int slot) {
    hipCheck(hipEventSynchronize(B.ev[slot][2]), "hipEventSynchronize");
    hipCheck(hipEventElapsedTime(&B.lastMs, B.ev[slot][0], B.ev[slot][1]), "hipEventElapsedTime");
    B.views = L.views;
    B.maxTiles = L.maxTiles;
    B.tilesPerView.resize(L.views);
    for (int v = 0; v < L.views; ++v) B.tilesPerView[v] = cps[v].tilesX * cps[v].tilesY;
    B.sortedGauss = B.sortedSlot[slot];
    B.lastCtl.assign(B.ctlHost[slot], B.ctlHost[slot] + L.views);
    uint32_t most = 0;
    bool over = false;
    for (const FrameCtl& c : B.lastCtl) {
        most = std::max(most, c.pairs);
        over |= c.overflow != 0;
    }
    if (over) {
        if (most >= (1u << 30) - 1) throw std::runtime_error("pair count of a view exceeds 2^30");
        B.pairCap = std::max(L.pairCap, static_cast<size_t>(most) + most / 4 + 1024);
        B.capByTile[tileKey(L.opts)] = B.pairCap;
    }
    return over;
}

bool waitBatchCounts(ViewBatch& B, const Layout& L, const std::vector<CamParams>& cps, int slot) {
    if (L.n == 0) return finishBatch(B, L, cps, slot);  // no pair-count scan ran
    for (uint32_t spin = 1;; ++spin) {
        int ready = 0;
        for (int v = 0; v < L.views; ++v) ready += __atomic_load_n(&B.mirror[v].seq, __ATOMIC_ACQUIRE) == 1u;
        if (ready == L.views) break;
        if ((spin & 1023) == 0) {
            // Has the batch ended without the counts arriving? Then its copied control blocks hold
            // them.
            const hipError_t q = hipEventQuery(B.ev[slot][2]);
            if (q == hipSuccess) return finishBatch(B, L, cps, slot);
            if (q != hipErrorNotReady) hipCheck(q, "hipEventQuery");
        }
        __builtin_ia32_pause();
    }
    bool over = false;
    for (int v = 0; v < L.views; ++v) over |= B.mirror[v].overflow != 0;
    if (over) {
        // The rest of the batch skips its work; let it end, then grow from its control blocks.
        hipCheck(hipEventSynchronize(B.ev[slot][2]), "hipEventSynchronize");
        return finishBatch(B, L, cps, slot);
    }
    B.lastMs = 0.0f;
    B.views = L.views;
    B.maxTiles = L.maxTiles;
    B.tilesPerView.resize(L.views);
    B.lastCtl.assign(L.views, FrameCtl{});
    for (int v = 0; v < L.views; ++v) {
        B.tilesPerView[v] = cps[v].tilesX * cps[v].tilesY;
        B.lastCtl[v].visible = B.mirror[v].visible;
        B.lastCtl[v].pairs = B.mirror[v].pairs;
    }
    B.sortedGauss = B.sortedSlot[slot];
    return false;
}

namespace {

ViewBatch& batchOf(ans_renderer* r) {
    if (!r->batch) r->batch.reset(new ViewBatch());
    return *r->batch;
}

// Validates a batch's arguments and derives each view's constants.
std::vector<CamParams> prepareViews(ans_renderer* r, int32_t n_views, const ans_camera* cams,
                                    const ans_options& o, uint8_t* const* dev_rgb) {
    if (!r) throw std::invalid_argument("renderer is null");
    if (n_views < 1 || n_views > kMaxViews) throw std::invalid_argument("n_views must be in 1..256");
    if (!cams || !dev_rgb) throw std::invalid_argument("cameras and outputs must not be null");
    std::vector<CamParams> cps(n_views);
    for (int v = 0; v < n_views; ++v) {
        checkCamera(&cams[v]);
        if (!dev_rgb[v]) throw std::invalid_argument("output pointer must not be null");
        cps[v] = makeCamParams(cams[v], o, r->scene.shDegree);
        if (cps[v].tilesX > 65535 || cps[v].tilesY > 65535) throw std::invalid_argument("too many tiles");
    }
    return cps;
}

// Ends the ans_render_views_async batch in `slot`, if one is in flight; true if it overflowed.
bool retireAsync(ViewBatch& B, int slot) {
    if (!B.asyncSeq[slot]) return false;
    B.asyncSeq[slot] = 0;
    return finishBatch(B, B.asyncL[slot], B.asyncCps[slot], slot);
}

// Ends every ans_render_views_async batch in flight, oldest first, and keeps any overflow for the
// next ans_views_wait.
void drainAsync(ViewBatch& B) {
    const int first = (B.asyncSeq[0] && (!B.asyncSeq[1] || B.asyncSeq[0] < B.asyncSeq[1])) ? 0 : 1;
    B.asyncOverflow |= retireAsync(B, first);
    B.asyncOverflow |= retireAsync(B, first ^ 1);
}

void renderViews(ans_renderer* r, int32_t n_views, const ans_camera* cams, const ans_options* options,
                 uint8_t* const* dev_rgb, ans_stats* stats, void* stream) {
    const ans_options o = resolveOptions(options);
    const std::vector<CamParams> cps = prepareViews(r, n_views, cams, o, dev_rgb);
    ViewBatch& B = batchOf(r);
    drainAsync(B);  // slot 0 and the arena are shared with asynchronous batches
    int retries = 0;
    while (true) {
        const Layout L = makeLayout(B, r->scene, n_views, cps, o);
        reserveArena(B, L);
        B.pairCap = L.pairCap;
        BatchIO io;
        io.outs = dev_rgb;
        submitBatch(B, r->scene, L, cps, io, static_cast<hipStream_t>(stream), 0);
        if (!finishBatch(B, L, cps, 0)) break;
        ++retries;
    }
    B.lastRetries = retries;
    if (stats) {
        for (int v = 0; v < n_views; ++v) {
            ans_stats& st = stats[v];
            st.visible = B.lastCtl[v].visible;
            st.tile_pairs = B.lastCtl[v].pairs;
            st.project_ms = 0.0f;
            st.sort_ms = 0.0f;
            st.blend_ms = B.lastMs;
            st.pair_capacity = static_cast<int64_t>(B.pairCap);
            st.retries = retries;
            st.reserved = 0;
        }
    }
}

// Enqueues one batch without waiting for it. Up to two batches are in flight, one per submission
// slot (its pinned table and control-block copy): a third call first ends the oldest. All of them
// share the arena, so a batch enqueued on another stream waits for the one in flight, and growth
// (a larger batch shape, or pairs above the per-view capacity) waits for every batch in flight.
void renderViewsAsync(ans_renderer* r, int32_t n_views, const ans_camera* cams, const ans_options* options,
                      uint8_t* const* dev_rgb, void* stream) {
    const ans_options o = resolveOptions(options);
    std::vector<CamParams> cps = prepareViews(r, n_views, cams, o, dev_rgb);
    ViewBatch& B = batchOf(r);
    const int slot = !B.asyncSeq[0] ? 0 : !B.asyncSeq[1] ? 1 : (B.asyncSeq[0] < B.asyncSeq[1] ? 0 : 1);
    B.asyncOverflow |= retireAsync(B, slot);
    Layout L = makeLayout(B, r->scene, n_views, cps, o);
    if (!arenaFits(B, L)) {
        drainAsync(B);  // growth frees buffers the batch in flight reads
        L = makeLayout(B, r->scene, n_views, cps, o);
    }
    reserveArena(B, L);
    B.pairCap = L.pairCap;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int other = slot ^ 1;
    if (B.asyncSeq[other]) hipCheck(hipStreamWaitEvent(st, B.ev[other][1], 0), "hipStreamWaitEvent");
    BatchIO io;
    io.outs = dev_rgb;
    submitBatch(B, r->scene, L, cps, io, st, slot);
    B.asyncSeq[slot] = B.asyncNext++;
    B.asyncViews = n_views;
    B.asyncL[slot] = L;
    B.asyncCps[slot] = std::move(cps);
}

// Waits for every asynchronous batch; true if one of them overflowed since the last wait.
bool waitViews(ans_renderer* r, ans_stats* stats) {
    if (!r) throw std::invalid_argument("renderer is null");
    ViewBatch& B = batchOf(r);
    drainAsync(B);
    const bool over = B.asyncOverflow;
    B.asyncOverflow = false;
    if (over) B.views = 0;  // a batch to render again leaves no lists to inspect
    if (stats)
        for (size_t v = 0; v < std::min(B.lastCtl.size(), static_cast<size_t>(B.asyncViews)); ++v) {
            ans_stats& st = stats[v];
            st.visible = B.lastCtl[v].visible;
            st.tile_pairs = B.lastCtl[v].pairs;
            st.project_ms = 0.0f;
            st.sort_ms = 0.0f;
            st.blend_ms = B.lastMs;
            st.pair_capacity = static_cast<int64_t>(B.pairCap);
            st.retries = 0;
            st.reserved = 0;
        }
    return over;
}

// One chunk of a camera path: frames [first, first + count), their constants, and each frame's
// byte offset in the chunk's output slot.
struct PathChunk {
    int64_t first = 0;
    int count = 0;
    std::vector<CamParams> cps;
    std::vector<size_t> offset;
    size_t bytes = 0;
};

void ensurePinned(uint8_t*& p, size_t& cap, size_t bytes) {
    if (bytes <= cap) return;
    if (p) hipCheck(hipHostFree(p), "hipHostFree");
    p = nullptr;
    cap = 0;
    hipCheck(hipHostMalloc(reinterpret_cast<void**>(&p), bytes, hipHostMallocDefault), "hipHostMalloc");
    cap = bytes;
}

// The frame loop of `aniso path` (src/main.cpp:234-244) as a two-deep pipeline of batched
// launches: chunk c + 1 renders on the path's render stream while chunk c is copied to pinned
// host memory on the copy stream and handed to the sink, frame by frame, in order.
void renderPath(ans_renderer* r, int64_t nFrames, const ans_camera* cams, const float* scales,
                const ans_options* options, int chunk, ans_frame_sink sink, void* user, ans_path_stats* stats) {
    const auto t0 = std::chrono::steady_clock::now();
    if (!r) throw std::invalid_argument("renderer is null");
    if (nFrames < 0) throw std::invalid_argument("n_frames must not be negative");
    if (chunk < 1 || chunk > kMaxViews) throw std::invalid_argument("chunk must be in 1..256");
    if (nFrames > 0 && (!cams || !sink)) throw std::invalid_argument("cameras and sink must not be null");
    const ans_options o = resolveOptions(options);

    // Every frame's constants, validated before any GPU work.
    std::vector<PathChunk> chunks;
    uint32_t maxTiles = 1;
    size_t maxBytes = 0;
    for (int64_t f = 0; f < nFrames; ++f) {
        checkCamera(&cams[f]);
        CamParams cp = makeCamParams(cams[f], o, r->scene.shDegree);
        if (scales) {
            if (!(scales[f] > 0.0f) || !std::isfinite(scales[f]))
                throw std::invalid_argument("splat scales must be positive and finite");
            cp.splatScale = scales[f];
        }
        if (cp.tilesX > 65535 || cp.tilesY > 65535) throw std::invalid_argument("too many tiles");
        if (chunks.empty() || chunks.back().count == chunk) {
            chunks.emplace_back();
            chunks.back().first = f;
        }
        PathChunk& k = chunks.back();
        k.cps.push_back(cp);
        k.offset.push_back(k.bytes);
        k.bytes += alignUp(static_cast<size_t>(cp.width) * cp.height * 3, 256);
        ++k.count;
        maxTiles = std::max(maxTiles, static_cast<uint32_t>(cp.tilesX) * static_cast<uint32_t>(cp.tilesY));
        maxBytes = std::max(maxBytes, k.bytes);
    }
    ans_path_stats st{};
    if (chunks.empty()) {
        if (stats) *stats = st;
        return;
    }

    ViewBatch& B = batchOf(r);
    drainAsync(B);  // the path's slots and the arena are shared with asynchronous batches
    if (!B.pathRender) {
        hipCheck(hipStreamCreateWithFlags(&B.pathRender, hipStreamNonBlocking), "hipStreamCreate");
        hipCheck(hipStreamCreateWithFlags(&B.pathCopy, hipStreamNonBlocking), "hipStreamCreate");
        for (auto& e : B.copied) hipCheck(hipEventCreate(&e), "hipEventCreate");
    }
    // The arena for the largest chunk shape up front: growth frees buffers, so it only happens
    // below after the render stream has drained.
    {
        Layout env = makeLayout(B, r->scene, chunks[0].count, chunks[0].cps, o);
        env.views = std::min<int>(chunk, static_cast<int>(nFrames));
        env.maxTiles = maxTiles;
        env.items = maxTiles * static_cast<uint32_t>(env.views);
        env.blocksI = std::max<uint32_t>(1, (env.items + kSortBlock * kDepthItems - 1) / (kSortBlock * kDepthItems));
        reserveArena(B, env);
        B.pairCap = env.pairCap;
    }
    for (int k = 0; k < 2; ++k) {
        B.pathDev[k].reserve(maxBytes);
        ensurePinned(B.pathHost[k], B.pathHostCap[k], maxBytes);
    }

    auto submit = [&](size_t c) {
        const PathChunk& k = chunks[c];
        const int slot = static_cast<int>(c & 1);
        const Layout L = makeLayout(B, r->scene, k.count, k.cps, o);
        reserveArena(B, L);
        B.pairCap = L.pairCap;
        std::vector<uint8_t*> outs(k.count);
        for (int v = 0; v < k.count; ++v) outs[v] = B.pathDev[slot].p + k.offset[v];
        BatchIO io;
        io.outs = outs.data();
        submitBatch(B, r->scene, L, k.cps, io, B.pathRender, slot);
        ++st.chunks;
        return L;
    };
    auto drain = [&] {
        (void)hipStreamSynchronize(B.pathRender);
        (void)hipStreamSynchronize(B.pathCopy);
    };
    try {
        Layout cur = submit(0);
        for (size_t c = 0; c < chunks.size(); ++c) {
            const PathChunk& k = chunks[c];
            const int slot = static_cast<int>(c & 1);
            // The next chunk's slot was last used by chunk c - 1, handed to the sink already.
            Layout next{};
            if (c + 1 < chunks.size()) next = submit(c + 1);
            bool over = finishBatch(B, cur, k.cps, slot);
            st.gpu_ms += B.lastMs;
            while (over) {
                // Pairs overflowed this chunk's arena: let the chunk in flight finish (growth frees
                // buffers), grow, and render this chunk again. The chunk in flight is checked on
                // its own turn and rendered again there if it overflowed too.
                hipCheck(hipStreamSynchronize(B.pathRender), "hipStreamSynchronize");
                const Layout L = submit(c);
                over = finishBatch(B, L, k.cps, slot);
                st.gpu_ms += B.lastMs;
            }
            hipCheck(hipMemcpyAsync(B.pathHost[slot], B.pathDev[slot].p, k.bytes, hipMemcpyDeviceToHost, B.pathCopy),
                     "hipMemcpyAsync frames");
            hipCheck(hipEventRecord(B.copied[slot], B.pathCopy), "hipEventRecord");
            hipCheck(hipEventSynchronize(B.copied[slot]), "hipEventSynchronize");
            for (int v = 0; v < k.count; ++v) {
                const CamParams& cp = k.cps[v];
                const int64_t f = k.first + v;
                const int rc = sink(user, f, B.pathHost[slot] + k.offset[v], static_cast<size_t>(cp.width) * cp.height * 3);
                if (rc != 0)
                    throw std::invalid_argument("sink returned " + std::to_string(rc) + " at frame " + std::to_string(f));
                ++st.frames;
            }
            cur = next;
        }
    } catch (...) {
        drain();
        B.views = 0;
        throw;
    }
    // A path leaves no batch to inspect: an arena growth while a chunk was in flight would leave
    // that chunk's lists in freed buffers, so the per-view counts and dumps refuse after a path.
    B.views = 0;
    st.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (stats) *stats = st;
}

}  // namespace

}  // namespace ans

using namespace ans;

extern "C" {

int ans_render_views(ans_renderer* r, int32_t n_views, const ans_camera* cams, const ans_options* options,
                     uint8_t* const* dev_rgb, void* stream, ans_stats* stats) {
    return guarded([&] {
        renderViews(r, n_views, cams, options, dev_rgb, stats, stream);
        return ANS_OK;
    });
}

int ans_render_views_async(ans_renderer* r, int32_t n_views, const ans_camera* cams, const ans_options* options,
                           uint8_t* const* dev_rgb, void* stream) {
    return guarded([&] {
        renderViewsAsync(r, n_views, cams, options, dev_rgb, stream);
        return ANS_OK;
    });
}

int ans_views_wait(ans_renderer* r, ans_stats* stats) {
    return guarded([&] { return waitViews(r, stats) ? ANS_ERR_CAPACITY : ANS_OK; });
}

int ans_views_autotune(ans_renderer* r, int32_t n_views, const ans_camera* cams, const ans_options* base,
                       uint8_t* const* dev_rgb, void* stream, int32_t repeats, ans_options* best, double* best_ms) {
    return guarded([&] {
        if (!best) throw std::invalid_argument("best is null");
        const ans_options b = resolveOptions(base);
        const int reps = std::max(1, repeats);
        static const int kTiles[3][2] = {{8, 8}, {16, 16}, {32, 16}};
        static const int kDepths[3] = {64, 128, 256};
        double bestMs = 1e30;
        ans_options bestOpt = b;
        for (const auto& t : kTiles) {
            for (int d : kDepths) {
                ans_options o = b;
                o.tile_w = t[0];
                o.tile_h = t[1];
                o.batch_depth = d;
                renderViews(r, n_views, cams, &o, dev_rgb, nullptr, stream);  // builds and warms the graph
                double ms = 1e30;
                for (int k = 0; k < reps; ++k) {
                    renderViews(r, n_views, cams, &o, dev_rgb, nullptr, stream);
                    ms = std::min(ms, static_cast<double>(r->batch->lastMs));
                }
                if (ms < bestMs) {
                    bestMs = ms;
                    bestOpt = o;
                }
            }
        }
        *best = bestOpt;
        if (best_ms) *best_ms = bestMs;
        return ANS_OK;
    });
}

int ans_render_path(ans_renderer* r, int64_t n_frames, const ans_camera* cams, const float* splat_scales,
                    const ans_options* options, int32_t chunk, ans_frame_sink sink, void* user,
                    ans_path_stats* stats) {
    return guarded([&] {
        renderPath(r, n_frames, cams, splat_scales, options, chunk, sink, user, stats);
        return ANS_OK;
    });
}

int ans_views_frame_counts(ans_renderer* r, int32_t view, int64_t* visible, int64_t* pairs, int64_t* tiles) {
    return guarded([&] {
        if (!r || !r->batch || r->batch->views == 0) throw std::invalid_argument("no batched render on this handle");
        ViewBatch& B = *r->batch;
        if (view < 0 || view >= B.views) throw std::invalid_argument("view out of range");
        if (visible) *visible = B.lastCtl[view].visible;
        if (pairs) *pairs = B.lastCtl[view].pairs;
        if (tiles) *tiles = B.tilesPerView[view];
        return ANS_OK;
    });
}

int ans_views_frame_dump(ans_renderer* r, int32_t view, uint32_t* values, uint32_t* ranges) {
    return guarded([&] {
        if (!r || !r->batch || r->batch->views == 0) throw std::invalid_argument("no batched render on this handle");
        ViewBatch& B = *r->batch;
        if (view < 0 || view >= B.views) throw std::invalid_argument("view out of range");
        const size_t P = B.lastCtl[view].pairs, T = static_cast<size_t>(B.tilesPerView[view]);
        if (values && P)
            hipCheck(hipMemcpy(values, B.sortedGauss + static_cast<size_t>(view) * B.pairCap, 4 * P,
                               hipMemcpyDeviceToHost),
                     "hipMemcpy values");
        if (ranges && T)
            hipCheck(hipMemcpy(ranges, B.ranges.p + static_cast<size_t>(view) * B.maxTiles, 8 * T,
                               hipMemcpyDeviceToHost),
                     "hipMemcpy ranges");
        return ANS_OK;
    });
}

}  // extern "C"
