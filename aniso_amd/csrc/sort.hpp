// Hand-written stable LSD radix sort pass for gfx950: one kernel per digit, onesweep style.
//
// Each block claims a partition of kSortTile consecutive items through an atomic counter, ranks
// its items by digit with wavefront ballots (stable: items keep their input order within a
// digit), publishes its per-digit counts, finds the counts of all earlier partitions by
// decoupled look-back, and scatters through LDS so that each digit's run is written to HBM
// with consecutive addresses. The global digit histogram comes from an earlier kernel.
//
// Keys and values are 32-bit. Replaces cub::DeviceRadixSort::SortPairs (gpu_render.cu:338-344)
// in a design that sorts 32-bit depth keys per Gaussian and 32-bit tile ids per pair instead of
// 64-bit tile|depth keys per pair.
#pragma once

#include "common.hpp"

namespace ans {

constexpr int kSortBlock = 256;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortBlock * kSortItems;
constexpr int kRadix = 256;
constexpr int kSortWaves = kSortBlock / kWave;

// Inclusive scan of one value per lane across a 64-lane wavefront.
__device__ inline uint32_t waveInclusiveScan(uint32_t v) {
    const int lane = __lane_id();
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const uint32_t t = __shfl_up(v, d, kWave);
        if (lane >= d) v += t;
    }
    return v;
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
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < MAXBITS; ++b) {
        if (b < bits) {
            const bool bit = (digit >> b) & 1u;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
    }
    return peers;
}

// One stable sort pass over bits [shift, shift + bits) of the keys (bits <= 8).
//   keysIn / valsIn: input (valsIn null: value = item index).
//   keysOut: may be null when the keys are not needed after this pass.
//   count: device item count (null: use nHost); items beyond cap are ignored.
//   hist: global histogram of this digit over all items (256 bins).
//   status: [parts][256] look-back words; tag = epoch * 16 + pass.
__global__ __launch_bounds__(kSortBlock) void k_sortPass(const uint32_t* __restrict__ keysIn,
                                                         const uint32_t* __restrict__ valsIn,
                                                         uint32_t* __restrict__ keysOut,
                                                         uint32_t* __restrict__ valsOut,
                                                         const uint32_t* __restrict__ count, uint32_t nHost,
                                                         uint32_t cap, int shift, int bits,
                                                         const uint32_t* __restrict__ hist,
                                                         uint32_t* __restrict__ partCounter,
                                                         uint64_t* __restrict__ status,
                                                         const FrameCtl* __restrict__ ctl, uint32_t pass) {
    __shared__ uint32_t sWaveHist[kSortWaves][kRadix];
    __shared__ uint32_t sBinStart[kRadix];
    __shared__ uint32_t sDst[kRadix];
    __shared__ uint32_t sKeys[kSortTile];
    __shared__ uint32_t sVals[kSortTile];
    __shared__ uint32_t sScan[kSortWaves];
    __shared__ uint32_t sPart;

    if (ctl->overflow) return;
    const uint32_t n = count ? min(*count, cap) : nHost;
    const int tid = threadIdx.x;
    const int lane = __lane_id();
    const int wave = tid / kWave;
    if (tid == 0) sPart = atomicAdd(partCounter, 1u);
    for (int i = tid; i < kSortWaves * kRadix; i += kSortBlock) (&sWaveHist[0][0])[i] = 0;
    __syncthreads();
    const uint32_t part = sPart;
    const uint32_t numParts = (n + kSortTile - 1) / kSortTile;
    if (part >= numParts) return;
    const uint32_t base = part * kSortTile;
    const uint32_t mask = (1u << bits) - 1u;

    // Load in wave-striped order: item j of lane l in wave w is base + (w * items + j) * 64 + l,
    // so (wave, j, lane) order is input order.
    uint32_t key[kSortItems], val[kSortItems], rank[kSortItems];
#pragma unroll
    for (int j = 0; j < kSortItems; ++j) {
        const uint32_t idx = base + (wave * kSortItems + j) * kWave + lane;
        const bool valid = idx < n;
        key[j] = valid ? keysIn[idx] : 0u;
        val[j] = valid ? (valsIn ? valsIn[idx] : idx) : 0u;
    }
    const uint64_t below = lanesBelow();
#pragma unroll
    for (int j = 0; j < kSortItems; ++j) {
        const uint32_t idx = base + (wave * kSortItems + j) * kWave + lane;
        const bool valid = idx < n;
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
    }
    __syncthreads();

    // Per digit (one thread per bin): exclusive offsets across waves, block totals, block-local
    // digit starts, then the look-back for the global start of this partition's run.
    const int b = tid;
    uint32_t run = 0;
#pragma unroll
    for (int w = 0; w < kSortWaves; ++w) {
        const uint32_t c = sWaveHist[w][b];
        sWaveHist[w][b] = run;
        run += c;
    }
    uint32_t blockTotal;
    const uint32_t binStart = blockExclusiveScan<kSortBlock>(run, sScan, &blockTotal);
    uint32_t histTotal;
    const uint32_t globalStart = blockExclusiveScan<kSortBlock>(hist[b], sScan, &histTotal);
    sBinStart[b] = binStart;
    const uint32_t tag = ctl->epoch * 16u + pass;
    uint64_t* st = status + static_cast<uint64_t>(part) * kRadix + b;
    uint32_t earlier = 0;
    if (part == 0) {
        storeStatus(st, statusWord(tag, kFlagPrefix, run));
    } else {
        storeStatus(st, statusWord(tag, kFlagAggregate, run));
        earlier = lookBack(status, part, kRadix, b, tag);
        storeStatus(st, statusWord(tag, kFlagPrefix, earlier + run));
    }
    sDst[b] = globalStart + earlier - binStart;
    __syncthreads();

    // Stage in block-sorted order, then write each digit's run contiguously.
#pragma unroll
    for (int j = 0; j < kSortItems; ++j) {
        const uint32_t idx = base + (wave * kSortItems + j) * kWave + lane;
        if (idx < n) {
            const uint32_t digit = (key[j] >> shift) & mask;
            const uint32_t pos = sBinStart[digit] + sWaveHist[wave][digit] + rank[j];
            sKeys[pos] = key[j];
            sVals[pos] = val[j];
        }
    }
    __syncthreads();
    const uint32_t inPart = min(static_cast<uint32_t>(kSortTile), n - base);
    for (uint32_t i = tid; i < inPart; i += kSortBlock) {
        const uint32_t k = sKeys[i];
        const uint32_t dst = sDst[(k >> shift) & mask] + i;
        if (keysOut) keysOut[dst] = k;
        valsOut[dst] = sVals[i];
    }
}

// Digit histograms of 32-bit keys, all four 8-bit digits at once, into hist[4][256].
__global__ __launch_bounds__(256) void k_histogram4(const uint32_t* __restrict__ keys, uint32_t n,
                                                    uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[4 * kRadix];
    for (int i = threadIdx.x; i < 4 * kRadix; i += 256) h[i] = 0;
    __syncthreads();
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const uint32_t k = keys[i];
#pragma unroll
        for (int d = 0; d < 4; ++d) atomicAdd(&h[d * kRadix + ((k >> (8 * d)) & 255u)], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 4 * kRadix; i += 256)
        if (h[i]) atomicAdd(&hist[i], h[i]);
}

}  // namespace ans
