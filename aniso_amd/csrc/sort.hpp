// Hand-written stable LSD radix sort for gfx950, reduce-then-scan. Per digit pass:
//   k_radixUpsweep    per-block digit histograms -> hist[bin][block] (one row per digit);
//   k_scanRows        one workgroup per digit scans its row in place and writes the row total;
//   k_radixDownsweep  re-reads its block, ranks items by digit with wavefront ballots (stable:
//                     equal digits keep input order), adds the digit's global start (a block
//                     scan of the row totals) and its row prefix, stages the items in LDS in
//                     digit order and writes each digit's run to consecutive HBM addresses.
// No block ever waits on another block, so there is no serial chain across the chip's XCDs.
//
// Keys and values are 32-bit. Replaces cub::DeviceRadixSort::SortPairs (gpu_render.cu:338-344)
// in a design that sorts 32-bit depth keys per Gaussian and 32-bit tile ids per pair instead of
// 64-bit tile|depth keys per pair.
//
// Batched views: blockIdx.y is the view. Each view sorts its own segment (items at
// view * itemStride, histograms at view * histStride, row totals at view * 256, control block
// ctl[view]); a single frame is the case gridDim.y = 1. With countFromCtl the item count is the
// view's pair count (ctl->pairs, at most cap) and blocks past it do no work.
#pragma once

#include "common.hpp"

namespace ans {

constexpr int kSortBlock = 256;
constexpr int kRadix = 256;
constexpr int kSortWaves = kSortBlock / kWave;
constexpr int kDepthItems = 8;  // 2048 keys per sort block (per-Gaussian sorts)
constexpr int kPairItems = 16;  // 4096 keys per sort block (the batched views' sorts)
constexpr int kTileItems = 8;   // 2048 keys per sort block (a single frame's tile sort)

// Inclusive scan of one value per lane across a 64-lane wavefront.
__device__ inline uint32_t waveInclusiveScan(uint32_t v) {
    return static_cast<uint32_t>(
        waveInclusive(static_cast<int>(v), 0, [](int a, int b) { return static_cast<int>(static_cast<uint32_t>(a) + static_cast<uint32_t>(b)); }));
}

// Exclusive scan across a block of BLOCK threads; sTmp holds BLOCK / 64 words. Returns the
// exclusive prefix and writes the block total to *total.
template <int BLOCK>
__device__ inline uint32_t blockExclusiveScan(uint32_t v, uint32_t* sTmp, uint32_t* total) {
    constexpr int waves = BLOCK / kWave;
    const int lane = __lane_id();
    const int wave = threadIdx.x / kWave;
    const uint32_t incl = waveInclusiveScan(v);
    if (lane == kWave - 1) sTmp[wave] = incl;
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < waves; ++w) {
        const uint32_t s = sTmp[w];
        before += (w < wave) ? s : 0;
        all += s;
    }
    __syncthreads();
    *total = all;
    return before + incl - v;
}

// Lanes of the wavefront whose digit equals this lane's (restricted to `valid` lanes).
template <int MAXBITS>
__device__ inline uint64_t matchDigit(uint32_t digit, int bits, bool valid) {
    uint64_t peers = __builtin_amdgcn_ballot_w64(valid);
#pragma unroll
    for (int b = 0; b < MAXBITS; ++b) {
        if (b < bits) {
            const bool bit = (digit >> b) & 1u;
            const uint64_t m = __builtin_amdgcn_ballot_w64(bit);
            peers &= bit ? m : ~m;
        }
    }
    return peers;
}

// Where a sort pass takes its item count (countFrom; a bool argument means host or pairs):
//   kCountHost    nHost;
//   kCountPairs   the frame's pair count (ctl->pairs, at most cap), or skip the pass if the pair
//                 buffers overflowed;
//   kCountKept    ctl->sortCount (at most nHost): the keys an earlier kCountDrop pass kept;
//   kCountDrop    nHost, dropping keys equal to 0xffffffff (culled Gaussians: they would sort
//                 last) — the pass writes only the kept keys, in order, and publishes their
//                 number in ctl->sortCount.
enum : int { kCountHost = 0, kCountPairs = 1, kCountKept = 2, kCountDrop = 3 };

// The number of items to sort, or -1 if this sort is skipped (the pair buffers overflowed). Only
// sorts that need it read the control block: a sort with a host-given count runs before anything
// of its frame can set the overflow flag, and a load it waited for would delay every other load
// of the workgroup by a round trip.
__device__ inline int64_t itemCount(const FrameCtl* ctl, int countFrom, uint32_t nHost, uint32_t cap) {
    if (countFrom == kCountHost || countFrom == kCountDrop) return nHost;
    if (countFrom == kCountKept) return min(ctl->sortCount, nHost);
    if (ctl->overflow) return -1;
    return min(ctl->pairs, cap);
}

// Per-block histogram of bits [shift, shift + bits) into hist[bin * stride + block], for the
// blocks that hold items.
template <int ITEMS>
__global__ __launch_bounds__(kSortBlock) void k_radixUpsweep(const uint32_t* __restrict__ keys, int countFrom,
                                                             uint32_t nHost, uint32_t cap, int shift, int bits,
                                                             uint32_t* __restrict__ hist, uint32_t stride,
                                                             const FrameCtl* __restrict__ ctl, uint32_t itemStride,
                                                             uint32_t histStride) {
    // kCopies histograms, lane l adding into copy l % kCopies: lanes of one LDS add that share a
    // digit (most of them when the digit is the top byte of nearby depths) update different words
    // instead of queueing on one. Rows are kRadix + 1 words apart so equal digits of different
    // copies fall in different banks.
    constexpr int kCopies = 8;
    constexpr int kRow = kRadix + 1;
    __shared__ uint32_t h[kCopies * kRow];
    const uint32_t view = blockIdx.y;
    ctl += view;
    const int64_t count = itemCount(ctl, countFrom, nHost, cap);
    if (count < 0) return;
    const uint32_t n = static_cast<uint32_t>(count);
    const bool drop = countFrom == kCountDrop;
    keys += static_cast<size_t>(view) * itemStride;
    hist += static_cast<size_t>(view) * histStride;
    const uint32_t base = blockIdx.x * (kSortBlock * ITEMS);
    if (base >= n && blockIdx.x > 0) return;
    const int tid = threadIdx.x;
    const int radix = 1 << bits;
    for (int i = tid; i < kCopies * kRow; i += kSortBlock) h[i] = 0;
    __syncthreads();
    const uint32_t mask = radix - 1;
    // Every key is loaded before any is counted (a load per LDS add would wait out each load's
    // latency in turn); past the end the index is clamped and the add skipped.
    if (base < n) {  // block 0 of an empty sort writes zeros
        uint32_t key[ITEMS];
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) key[j] = keys[min(base + j * kSortBlock + tid, n - 1)];
        uint32_t* mine = h + (__lane_id() % kCopies) * kRow;
#pragma unroll
        for (int j = 0; j < ITEMS; ++j)
            if (base + j * kSortBlock + tid < n && !(drop && key[j] == 0xffffffffu))
                atomicAdd(&mine[(key[j] >> shift) & mask], 1u);
    }
    __syncthreads();
    for (int i = tid; i < radix; i += kSortBlock) {
        uint32_t c = 0;
#pragma unroll
        for (int k = 0; k < kCopies; ++k) c += h[k * kRow + i];
        hist[static_cast<size_t>(i) * stride + blockIdx.x] = c;
    }
}

// Exclusive scan of one row hist[bin * stride, + blocks holding items) in place by one workgroup
// (blockIdx.x = bin); the row total goes to rowTotal[bin]. tileItems: items per sort block.
[[maybe_unused]] static __global__ __launch_bounds__(kSortBlock) void k_scanRows(uint32_t* __restrict__ hist, uint32_t stride,
                                                                uint32_t len, uint32_t* __restrict__ rowTotal,
                                                                const FrameCtl* __restrict__ ctl, int countFrom,
                                                                uint32_t cap, uint32_t tileItems,
                                                                uint32_t histStride) {
    __shared__ uint32_t sTmp[kSortWaves];
    const uint32_t view = blockIdx.y;
    ctl += view;
    const int64_t count = itemCount(ctl, countFrom, cap, cap);
    if (count < 0) return;
    hist += static_cast<size_t>(view) * histStride;
    rowTotal += view * kRadix;
    if (countFrom == kCountPairs || countFrom == kCountKept)
        len = max(1u, min(len, (static_cast<uint32_t>(count) + tileItems - 1) / tileItems));
    constexpr int kPer = 8;
    uint32_t* row = hist + static_cast<size_t>(blockIdx.x) * stride;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < len; base += kSortBlock * kPer) {
        const uint32_t first = base + threadIdx.x * kPer;
        uint32_t v[kPer];
        uint32_t sum = 0;
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            v[k] = (first + k < len) ? row[first + k] : 0u;
            sum += v[k];
        }
        uint32_t total;
        uint32_t run = carry + blockExclusiveScan<kSortBlock>(sum, sTmp, &total);
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            if (first + k < len) row[first + k] = run;
            run += v[k];
        }
        carry += total;
    }
    if (threadIdx.x == 0) rowTotal[blockIdx.x] = carry;
}

// One stable pass over bits [shift, shift + bits) (bits <= 8). keysOut may be null when the keys
// are not needed after this pass; valsIn null means value = item index. hist holds the
// row-scanned histogram of this pass and rowTotal the per-digit totals.
template <int ITEMS>
__global__ __launch_bounds__(kSortBlock) void k_radixDownsweep(
    const uint32_t* __restrict__ keysIn, const uint32_t* __restrict__ valsIn, uint32_t* __restrict__ keysOut,
    uint32_t* __restrict__ valsOut, int countFrom, uint32_t nHost, uint32_t cap, int shift, int bits,
    const uint32_t* __restrict__ hist, const uint32_t* __restrict__ rowTotal, uint32_t stride,
    FrameCtl* __restrict__ ctl, uint32_t itemStride, uint32_t histStride) {
    constexpr int kTile = kSortBlock * ITEMS;
    __shared__ uint32_t sWaveHist[kSortWaves][kRadix];
    __shared__ uint32_t sBinStart[kRadix];
    __shared__ uint32_t sDst[kRadix];
    __shared__ uint32_t sKeys[kTile];
    __shared__ uint32_t sVals[kTile];
    __shared__ uint32_t sScan[kSortWaves];

    const uint32_t view = blockIdx.y;
    const int tid = threadIdx.x;
    const uint32_t mask = (1u << bits) - 1u;
    // This thread's digit's total and row prefix, loaded first (before the count) so that their
    // latency overlaps the count's and the key loads' instead of following them.
    const bool live = tid <= static_cast<int>(mask);
    uint32_t digitTotal = 0, rowPrefix = 0;
    if (live) {
        digitTotal = rowTotal[view * kRadix + tid];
        rowPrefix = hist[static_cast<size_t>(view) * histStride + static_cast<size_t>(tid) * stride + blockIdx.x];
    }
    ctl += view;
    const int64_t count = itemCount(ctl, countFrom, nHost, cap);
    if (count < 0) return;
    const uint32_t n = static_cast<uint32_t>(count);
    const bool drop = countFrom == kCountDrop;
    const uint32_t base = blockIdx.x * kTile;
    if (base >= n) return;
    {
        const size_t io = static_cast<size_t>(view) * itemStride;
        keysIn += io;
        if (valsIn) valsIn += io;
        if (keysOut) keysOut += io;
        valsOut += io;
    }
    const int lane = __lane_id();
    const int wave = tid / kWave;
    for (int i = tid; i < kSortWaves * kRadix; i += kSortBlock) (&sWaveHist[0][0])[i] = 0;
    __syncthreads();

    // Wave-striped order: item j of lane l in wave w is base + (w * ITEMS + j) * 64 + l, so
    // (wave, j, lane) order is input order.
    uint32_t key[ITEMS], val[ITEMS], rank[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const uint32_t idx = base + (wave * ITEMS + j) * kWave + lane;
        const bool valid = idx < n;
        key[j] = valid ? keysIn[idx] : 0u;
        val[j] = valid ? (valsIn ? valsIn[idx] : idx) : 0u;
    }
    const uint64_t below = lanesBelow();
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const uint32_t idx = base + (wave * ITEMS + j) * kWave + lane;
        const bool valid = idx < n && !(drop && key[j] == 0xffffffffu);
        const uint32_t digit = (key[j] >> shift) & mask;
        const uint64_t peers = matchDigit<8>(digit, bits, valid);
        uint32_t r = 0;
        if (valid) {
            const uint32_t before = __popcll(peers & below);
            const uint32_t old = sWaveHist[wave][digit];
            r = old + before;
            if (before == 0) sWaveHist[wave][digit] = old + static_cast<uint32_t>(__popcll(peers));
        }
        rank[j] = r;
        waveSync();
    }
    __syncthreads();

    // One thread per digit: offsets across waves, block-local digit start, global destination.
    const int b = tid;
    uint32_t run = 0;
#pragma unroll
    for (int w = 0; w < kSortWaves; ++w) {
        const uint32_t c = sWaveHist[w][b];
        sWaveHist[w][b] = run;
        run += c;
    }
    uint32_t blockTotal, allTotal;
    const uint32_t binStart = blockExclusiveScan<kSortBlock>(run, sScan, &blockTotal);
    const uint32_t digitStart = blockExclusiveScan<kSortBlock>(digitTotal, sScan, &allTotal);
    sBinStart[b] = binStart;
    if (live) sDst[b] = digitStart + rowPrefix - binStart;
    if (drop && blockIdx.x == 0 && tid == 0) ctl[0].sortCount = allTotal;  // ctl is this view's
    __syncthreads();

#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const uint32_t idx = base + (wave * ITEMS + j) * kWave + lane;
        if (idx < n && !(drop && key[j] == 0xffffffffu)) {
            const uint32_t digit = (key[j] >> shift) & mask;
            const uint32_t pos = sBinStart[digit] + sWaveHist[wave][digit] + rank[j];
            sKeys[pos] = key[j];
            sVals[pos] = val[j];
        }
    }
    __syncthreads();
    const uint32_t inBlock = blockTotal;  // the block's items, less any dropped
    for (uint32_t i = tid; i < inBlock; i += kSortBlock) {
        const uint32_t k = sKeys[i];
        const uint32_t dst = sDst[(k >> shift) & mask] + i;
        if (keysOut) keysOut[dst] = k;
        valsOut[dst] = sVals[i];
    }
}

}  // namespace ans
