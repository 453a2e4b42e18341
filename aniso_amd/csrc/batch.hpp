// Host side of the batched launches (many cameras of one scene in one sequence of launches,
// replayed as a captured HIP graph): the per-call device table, the grow-only arena and the
// launch sequence shared by the batched render (batch.hip) and the batched differentiable frames
// (diffviews.hip).
#pragma once

#include "engine.hpp"
#include "views.hpp"

#include <map>
#include <tuple>
#include <vector>

namespace ans {

constexpr int kMaxViews = 256;

inline size_t alignUp(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Byte offsets of the per-call device table: CamParams[V], RGB8 outputs uint8_t*[V], float
// outputs float*[V], first pixel of each view in the pixel-state array uint64_t[V], tile-grid
// widths int[V], tile counts int[V], image gradients const float*[V], first pair of each view in
// the partial-gradient array uint64_t[V], then the parameter and gradient arrays
// (DiffParamsDev). Kernels read everything that changes between calls from here, so a captured
// graph replays for any poses, outputs and tensors.
struct TableLayout {
    size_t cams, outs, outf, pixOff, tilesX, tiles, grads, pairBase, params, bytes;
    explicit TableLayout(int v) {
        cams = 0;
        outs = alignUp(sizeof(CamParams) * v, 16);
        outf = alignUp(outs + sizeof(uint8_t*) * v, 16);
        pixOff = alignUp(outf + sizeof(float*) * v, 16);
        tilesX = alignUp(pixOff + sizeof(uint64_t) * v, 16);
        tiles = alignUp(tilesX + sizeof(int) * v, 16);
        grads = alignUp(tiles + sizeof(int) * v, 16);
        pairBase = alignUp(grads + sizeof(float*) * v, 16);
        params = alignUp(pairBase + sizeof(uint64_t) * v, 16);
        bytes = alignUp(params + sizeof(DiffParamsDev), 16);
    }
};

// Shapes of one batched call.
struct Layout {
    int views;
    uint32_t n;
    ans_options opts;
    bool cull;
    bool diff;        // differentiable: raw parameters, float images and pixel state
    uint64_t pixels;  // all views' pixels (differentiable batches)
    uint32_t maxTiles, totalTiles, items;
    int tileBits, passes, digitBits;
    uint32_t blocksD, blocksDup, blocksP, blocksI;
    size_t pairCap;
};

struct ViewBatch {
    // Grow-only arena. Cached graphs bake its addresses, so growth of any buffer drops them.
    DevBuf<uint8_t> table;
    DevBuf<FrameCtl> ctl;  // one per view, plus one for the item-order sort
    DevBuf<SplatRec> rec;
    DevBuf<uint32_t> depthKey, order, kA, vA, kB, vB, ttByRank;
    DevBuf<uint2> rectIdx, rectByRank;
    DevBuf<uint32_t> hist, rowTotal, blockPairs, blockVisible;
    DevBuf<uint32_t> tileA, gaussA, tileB, gaussB;
    DevBuf<uint2> ranges;
    DevBuf<uint32_t> itemKey, itemK2, itemV2, itemOrder;
    // Differentiable batches only: tile counts per (view, Gaussian) and per-pixel state
    // (final transmittance bits, last blended list position + 1) of every view.
    DevBuf<uint32_t> tilesTouched;
    DevBuf<uint2> pixState;

    hipStream_t stream = nullptr;  // graphs are captured here and launched on the caller's stream
    // Per submission slot (two batches may be in flight along a camera path): start and end of
    // the launches, control-block copy done; pinned staging of the per-view table; pinned copy of
    // the control blocks; the buffer the batch's sorted lists ended in.
    hipEvent_t ev[2][3] = {};
    // Differentiable batches: each view's pair count, published by the pair-count scan to pinned
    // host memory (seq 1 once written), so the host learns of an overflow without waiting for
    // the rest of the batch.
    CountMirror* mirror = nullptr;
    CountMirror* mirrorDev = nullptr;
    uint8_t* tableHost[2] = {};
    FrameCtl* ctlHost[2] = {};
    const uint32_t* sortedSlot[2] = {};
    size_t pairCap = 0;       // pairs per view of the last batch
    // Pairs per view by tile shape (tile_w << 16 | tile_h): a batch's capacity follows what its
    // own tile shape has needed, so that the sort grids of a 16x16 batch are not sized for the
    // pairs an 8x8 batch of the same views once had (the autotuner renders both).
    std::map<uint32_t, size_t> capByTile;
    uint64_t generation = 0;  // bumped when the arena moves

    // views, n, tile w, tile h, batch depth, cull, max tiles, total tiles, pair capacity,
    // generation, differentiable, total pixels
    using GraphKey =
        std::tuple<int, int64_t, int, int, int, bool, uint32_t, uint32_t, size_t, uint64_t, bool, uint64_t>;
    std::map<GraphKey, hipGraphExec_t> graphs;

    // Batches of ans_render_views_async not yet waited for, per submission slot: the order they
    // were submitted in (0: slot free), their shapes and cameras; and whether a batch finished
    // since the last wait overflowed the arena.
    uint64_t asyncSeq[2] = {};
    uint64_t asyncNext = 1;
    Layout asyncL[2] = {};
    std::vector<CamParams> asyncCps[2];
    bool asyncOverflow = false;
    int asyncViews = 0;  // views of the last batch enqueued asynchronously

    // The last finished batch, for stats and dumps.
    int views = 0;
    uint32_t maxTiles = 0;
    std::vector<int> tilesPerView;
    std::vector<FrameCtl> lastCtl;
    const uint32_t* sortedGauss = nullptr;
    float lastMs = 0.0f;
    int lastRetries = 0;

    // Camera-path streaming: its render and copy streams, per-slot device frames and pinned host
    // frames, copy-done events. Created on first use.
    hipStream_t pathRender = nullptr, pathCopy = nullptr;
    hipEvent_t copied[2] = {};
    DevBuf<uint8_t> pathDev[2];
    uint8_t* pathHost[2] = {};
    size_t pathHostCap[2] = {};

    ViewBatch();
    ~ViewBatch();
    ViewBatch(const ViewBatch&) = delete;
    ViewBatch& operator=(const ViewBatch&) = delete;
    void dropGraphs();
};

// What one batched call writes and reads besides the scene: per view an RGB8 output (render) or a
// float output (differentiable), and for differentiable calls the parameter arrays.
struct BatchIO {
    uint8_t* const* outs = nullptr;
    float* const* outf = nullptr;
    const DiffParamsDev* params = nullptr;
};

Layout makeLayout(const ViewBatch& B, const SceneDev& s, int views, const std::vector<CamParams>& cps,
                  const ans_options& o, bool diff = false);
void reserveArena(ViewBatch& B, const Layout& L);
// Whether the arena already holds the batch L (reserveArena would not move any buffer).
bool arenaFits(ViewBatch& B, const Layout& L);
// Enqueues the whole batch on st (kernel launches only, so it can be captured). Returns the
// buffer holding the per-view sorted Gaussian lists.
const uint32_t* enqueueBatch(ViewBatch& B, const SceneDev& s, const Layout& L, hipStream_t st);
// Uploads the per-view table (through submission slot `slot`'s pinned staging) and enqueues the
// batch on st: graph replay unless ANS_FLAG_NO_GRAPH, then the control blocks' copy to the slot's
// pinned buffer. Returns without waiting.
void submitBatch(ViewBatch& B, const SceneDev& s, const Layout& L, const std::vector<CamParams>& cps,
                 const BatchIO& io, hipStream_t st, int slot);
// Waits for the batch submitted through `slot`, makes it the last finished batch and returns true
// if any view's pairs overflowed the arena; the per-view capacity has then been raised.
bool finishBatch(ViewBatch& B, const Layout& L, const std::vector<CamParams>& cps, int slot);
// Differentiable batches: waits only until every view's pair count is published (the rest of
// the batch keeps running) and makes the batch the last one; on an overflow it waits for the
// batch and grows as finishBatch does, returning true. lastMs is left at 0.
bool waitBatchCounts(ViewBatch& B, const Layout& L, const std::vector<CamParams>& cps, int slot);

}  // namespace ans
