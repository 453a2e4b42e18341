// The differentiable path's handle: single-view frames (diff.hip) and batched multi-view frames
// (diffviews.hip) kept for their backward pass, and the scratch their backward passes share.
#pragma once

#include "batch.hpp"

#include <map>
#include <memory>
#include <vector>

namespace ans {

// The scene a differentiable call reads in place: the optimizer's tensors (ans_params).
SceneDev sceneOfParams(const ans_params* p);

// A batched differentiable forward kept for its backward: its own arena (records, lists, ranges,
// per-pixel state of every view), the shapes it was rendered with, and the captured backward
// graphs.
struct DiffViewsFrame {
    ViewBatch B;
    Layout L{};
    std::vector<CamParams> cps;
    int64_t n = 0;
    int shDegree = 0;
    bool act = false;
    hipEvent_t staged = nullptr;  // the backward's table upload has left the pinned staging
    // views, n, tile w, batch depth, max tiles, total tiles, pair capacity, arena generation,
    // scratch generation, total pixels
    using BwdKey = std::tuple<int, int64_t, int, int, uint32_t, uint32_t, size_t, uint64_t, uint64_t, uint64_t>;
    std::map<BwdKey, hipGraphExec_t> bwdGraphs;
    DiffViewsFrame();
    ~DiffViewsFrame();
    void dropBwdGraphs();
};

// Backward scratch shared by a handle's batched frames (one backward runs at a time per handle):
// first pair slot per (view, Gaussian) in index order, per-block sums for that scan, and the
// per-pair partial gradient records of all views, view v from pair pairBase[v].
struct DiffViewsScratch {
    DevBuf<uint32_t> slotStart, blockSums;
    DevBuf<float> partial;
    uint64_t generation = 0;
};

}  // namespace ans

struct ans_diff {
    std::unique_ptr<ans::Engine> engine;
    std::map<int64_t, std::unique_ptr<ans::FrameState>> frames;
    std::vector<std::unique_ptr<ans::FrameState>> pool;
    std::map<int64_t, std::unique_ptr<ans::DiffViewsFrame>> viewFrames;
    std::vector<std::unique_ptr<ans::DiffViewsFrame>> viewPool;
    ans::DiffViewsScratch viewScratch;
    int64_t next = 1;
};
