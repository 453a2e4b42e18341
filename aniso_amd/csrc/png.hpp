// PNG file images of RGB8 frames in the reference writer's exact byte layout (image.cpp:41-85):
// signature, IHDR (8-bit truecolour, no interlace), one IDAT holding a zlib stream of stored
// (uncompressed) deflate blocks of at most 65535 bytes over the scanlines (each prefixed with
// filter byte 0), the stream's Adler-32, and IEND; every chunk closed by its CRC-32.
//
// Two encoders with one layout: on the host (encodeHost) and on the device (enqueuePng, io.hip), where
// the frame is already in HBM. The device encoder copies the scanlines into their blocks with one
// thread per byte, reduces the Adler-32 sums per workgroup, and computes the IDAT CRC-32 in
// 256-byte pieces whose registers are combined in a tree: the CRC register is affine in the
// message, R(A || B) = Z^|B| R(A) xor R0(B), where R0 starts from zero and Z^k, the effect of k
// zero bytes, is a 32 x 32 matrix over GF(2) (the combination zlib's crc32_combine publishes).
// The host builds Z^(2^i) for i < 40 (constant memory) and a shift by k multiplies the matrices
// of the set bits of k.
#pragma once

#include "common.hpp"

#include <array>
#include <cstring>
#include <stdexcept>

namespace ans {
namespace png {

constexpr uint32_t kCrcPoly = 0xEDB88320u;  // CRC-32 (ISO 3309 / ITU-T V.42), reflected
constexpr uint32_t kAdlerMod = 65521u;      // largest prime below 2^16 (RFC 1950)
constexpr size_t kStoredMax = 65535;        // bytes per stored deflate block
constexpr int kCrcPiece = 256;              // bytes per CRC piece of the device encoder
constexpr int kShiftBits = 40;              // Z^(2^i), i < 40: lengths below 1 TB

// Layout of one file image for a w x h frame.
struct Layout {
    size_t stride;    // scanline bytes: filter byte + 3 w
    size_t raw;       // h * stride
    size_t blocks;    // stored blocks (one empty block when raw is 0)
    size_t zlib;      // 2 + 5 blocks + raw + 4
    size_t idatData;  // offset of the zlib stream in the file
    size_t total;     // file size
    explicit Layout(int w, int h) {
        stride = 1 + 3 * static_cast<size_t>(w);
        raw = static_cast<size_t>(h) * stride;
        blocks = raw == 0 ? 1 : (raw + kStoredMax - 1) / kStoredMax;
        zlib = 2 + 5 * blocks + raw + 4;
        idatData = 8 + 25 + 8;
        total = idatData + zlib + 4 + 12;
    }
};

// Byte-at-a-time table (entry b: the register after shifting byte b through eight steps).
inline const std::array<uint32_t, 256>& crcTable() {
    static const std::array<uint32_t, 256> t = [] {
        std::array<uint32_t, 256> r{};
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t c = i;
            for (int k = 0; k < 8; ++k) c = (c & 1u) ? kCrcPoly ^ (c >> 1) : c >> 1;
            r[i] = c;
        }
        return r;
    }();
    return t;
}

// Running CRC register over bytes (no pre- or post-inversion).
inline uint32_t crcRegister(uint32_t reg, const uint8_t* p, size_t n) {
    const auto& t = crcTable();
    for (size_t i = 0; i < n; ++i) reg = t[(reg ^ p[i]) & 0xffu] ^ (reg >> 8);
    return reg;
}

inline uint32_t crc32(const uint8_t* p, size_t n) { return ~crcRegister(0xffffffffu, p, n); }

// Adler-32 (RFC 1950) with the sums reduced only every 5552 bytes, the longest run for which
// b stays below 2^32 (the bound zlib's NMAX states).
inline uint32_t adler32(const uint8_t* p, size_t n) {
    uint32_t a = 1, b = 0;
    while (n > 0) {
        const size_t run = n < 5552 ? n : 5552;
        for (size_t i = 0; i < run; ++i) {
            a += p[i];
            b += a;
        }
        a %= kAdlerMod;
        b %= kAdlerMod;
        p += run;
        n -= run;
    }
    return (b << 16) | a;
}

inline void put32(uint8_t* o, uint32_t v) {
    o[0] = static_cast<uint8_t>(v >> 24);
    o[1] = static_cast<uint8_t>(v >> 16);
    o[2] = static_cast<uint8_t>(v >> 8);
    o[3] = static_cast<uint8_t>(v);
}

// The 41 bytes before the zlib stream (signature, IHDR, IDAT length and type) and the 12 bytes
// of IEND.
struct Frame {
    uint8_t head[41];
    uint8_t tail[12];
};

inline Frame frameBytes(int w, int h, const Layout& L) {
    Frame f{};
    static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    std::memcpy(f.head, sig, 8);
    uint8_t* ihdr = f.head + 8;
    put32(ihdr, 13);
    std::memcpy(ihdr + 4, "IHDR", 4);
    put32(ihdr + 8, static_cast<uint32_t>(w));
    put32(ihdr + 12, static_cast<uint32_t>(h));
    const uint8_t rest[5] = {8, 2, 0, 0, 0};  // bit depth 8, truecolour, deflate, filter 0, no interlace
    std::memcpy(ihdr + 16, rest, 5);
    put32(ihdr + 21, crc32(ihdr + 4, 17));
    put32(f.head + 33, static_cast<uint32_t>(L.zlib));
    std::memcpy(f.head + 37, "IDAT", 4);
    put32(f.tail, 0);
    std::memcpy(f.tail + 4, "IEND", 4);
    put32(f.tail + 8, crc32(f.tail + 4, 4));
    return f;
}

// Stored-block header of block k (5 bytes: BFINAL, LEN, NLEN little-endian).
inline void blockHeader(uint8_t* o, size_t k, const Layout& L) {
    const size_t len = L.raw == 0 ? 0 : std::min(kStoredMax, L.raw - k * kStoredMax);
    o[0] = k + 1 == L.blocks ? 1 : 0;
    o[1] = static_cast<uint8_t>(len & 0xffu);
    o[2] = static_cast<uint8_t>(len >> 8);
    o[3] = static_cast<uint8_t>(~len & 0xffu);
    o[4] = static_cast<uint8_t>((~len >> 8) & 0xffu);
}

// The whole file image on the host into out[L.total].
inline void encodeHost(const uint8_t* rgb, int w, int h, uint8_t* out) {
    const Layout L(w, h);
    const Frame f = frameBytes(w, h, L);
    std::memcpy(out, f.head, 41);
    uint8_t* z = out + L.idatData;
    z[0] = 0x78;
    z[1] = 0x01;
    uint32_t a = 1, b = 0;
    size_t since = 0;  // bytes added since the sums were last reduced
    uint8_t* o = z + 2;
    for (size_t k = 0; k < L.blocks; ++k) {
        blockHeader(o, k, L);
        o += 5;
        const size_t first = k * kStoredMax, last = std::min(L.raw, first + kStoredMax);
        for (size_t r = first; r < last; ++r) {
            const size_t row = r / L.stride, c = r - row * L.stride;
            const uint8_t v = c == 0 ? 0 : rgb[row * (L.stride - 1) + c - 1];
            *o++ = v;
            a += v;
            b += a;
            if (++since == 5552) {
                a %= kAdlerMod;
                b %= kAdlerMod;
                since = 0;
            }
        }
    }
    a %= kAdlerMod;
    b %= kAdlerMod;
    put32(o, (b << 16) | a);
    put32(z + L.zlib, crc32(out + L.idatData - 4, L.zlib + 4));
    std::memcpy(z + L.zlib + 4, f.tail, 12);
}

// Z^(2^i) for i < kShiftBits, column j of matrix i at [i][j]: the register bit j turns into
// after 2^i zero bytes.
struct ShiftTable {
    uint32_t m[kShiftBits][32];
};

inline uint32_t gfApply(const uint32_t* mat, uint32_t v) {
    uint32_t r = 0;
    for (int j = 0; v; ++j, v >>= 1)
        if (v & 1u) r ^= mat[j];
    return r;
}

inline const ShiftTable& shiftTable() {
    static const ShiftTable t = [] {
        ShiftTable s{};
        // One zero byte: eight single-bit steps of the reflected register.
        uint32_t one[32];
        for (int j = 0; j < 32; ++j) {
            uint32_t c = 1u << j;
            for (int k = 0; k < 8; ++k) c = (c & 1u) ? kCrcPoly ^ (c >> 1) : c >> 1;
            one[j] = c;
        }
        std::memcpy(s.m[0], one, sizeof(one));
        for (int i = 1; i < kShiftBits; ++i)
            for (int j = 0; j < 32; ++j) s.m[i][j] = gfApply(s.m[i - 1], s.m[i - 1][j]);
        return s;
    }();
    return t;
}

// ---- device encoder -----------------------------------------------------------------------

constexpr int kBodyBlock = 256;
constexpr int kBodyItems = 16;  // raw bytes per thread: 4096 per workgroup

struct DevLayout {
    uint64_t stride, raw, blocks, zlib, idatData;
};

// Scanline bytes into their stored blocks, block headers, and per-workgroup Adler-32 sums:
// part[2 wg] = sum of bytes, part[2 wg + 1] = sum of (raw - r) byte_r, both mod 65521.
__global__ __launch_bounds__(kBodyBlock) void k_pngBody(const uint8_t* __restrict__ rgb, uint8_t* __restrict__ out,
                                                        DevLayout L, uint32_t* __restrict__ part) {
    __shared__ uint64_t sA[kBodyBlock / kWave], sB[kBodyBlock / kWave];
    const uint64_t base = static_cast<uint64_t>(blockIdx.x) * (kBodyBlock * kBodyItems);
    uint8_t* z = out + L.idatData;
    uint64_t a = 0, b = 0;
    for (int j = 0; j < kBodyItems; ++j) {
        const uint64_t r = base + static_cast<uint64_t>(j) * kBodyBlock + threadIdx.x;
        if (r >= L.raw) break;
        const uint64_t row = r / L.stride, c = r - row * L.stride;
        const uint8_t v = c == 0 ? 0 : rgb[row * (L.stride - 1) + c - 1];
        const uint64_t k = r / kStoredMax;
        z[2 + 5 * (k + 1) + r] = v;
        a += v;
        b += (L.raw - r) * v;
        if (r == k * kStoredMax) {  // first byte of block k writes its header
            const uint64_t len = min(static_cast<uint64_t>(kStoredMax), L.raw - r);
            uint8_t* hdr = z + 2 + 5 * k + r;
            hdr[0] = k + 1 == L.blocks ? 1 : 0;
            hdr[1] = static_cast<uint8_t>(len & 0xffu);
            hdr[2] = static_cast<uint8_t>(len >> 8);
            hdr[3] = static_cast<uint8_t>(~len & 0xffu);
            hdr[4] = static_cast<uint8_t>((~len >> 8) & 0xffu);
        }
    }
    // Per-thread sums are below 2^8 * 16 and 2^40 * 2^4 * 2^8: no overflow in 64 bits.
    for (int d = kWave / 2; d > 0; d >>= 1) {
        a += __shfl_down(a, d, kWave);
        b += __shfl_down(b, d, kWave);
    }
    const int wave = threadIdx.x / kWave;
    if (__lane_id() == 0) {
        sA[wave] = a % kAdlerMod;
        sB[wave] = b % kAdlerMod;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t ta = 0, tb = 0;
        for (int w = 0; w < kBodyBlock / kWave; ++w) ta += sA[w], tb += sB[w];
        part[2 * blockIdx.x] = static_cast<uint32_t>(ta % kAdlerMod);
        part[2 * blockIdx.x + 1] = static_cast<uint32_t>(tb % kAdlerMod);
    }
}

// One workgroup: the file's fixed bytes (signature, IHDR, IDAT header, zlib header, IEND) and the
// Adler-32 from the partial sums: a = 1 + sum of bytes, b = raw + sum of (raw - r) byte_r.
__global__ __launch_bounds__(256) void k_pngAdler(uint8_t* __restrict__ out, DevLayout L, Frame f,
                                                  const uint32_t* __restrict__ part, uint32_t parts) {
    __shared__ uint64_t sA[256 / kWave], sB[256 / kWave];
    uint64_t a = 0, b = 0;
    for (uint32_t i = threadIdx.x; i < parts; i += 256) {
        a += part[2 * i];
        b += part[2 * i + 1];
    }
    for (int d = kWave / 2; d > 0; d >>= 1) {
        a += __shfl_down(a, d, kWave);
        b += __shfl_down(b, d, kWave);
    }
    const int wave = threadIdx.x / kWave;
    if (__lane_id() == 0) sA[wave] = a, sB[wave] = b;
    __syncthreads();
    if (threadIdx.x < 41) out[threadIdx.x] = f.head[threadIdx.x];
    uint8_t* z = out + L.idatData;
    if (threadIdx.x < 12) z[L.zlib + 4 + threadIdx.x] = f.tail[threadIdx.x];
    if (threadIdx.x == 0) {
        uint64_t ta = 1, tb = L.raw % kAdlerMod;
        for (int w = 0; w < 256 / kWave; ++w) ta += sA[w], tb += sB[w];
        const uint32_t adler = static_cast<uint32_t>(((tb % kAdlerMod) << 16) | (ta % kAdlerMod));
        z[0] = 0x78;
        z[1] = 0x01;
        uint8_t* o = z + L.zlib - 4;
        o[0] = static_cast<uint8_t>(adler >> 24);
        o[1] = static_cast<uint8_t>(adler >> 16);
        o[2] = static_cast<uint8_t>(adler >> 8);
        o[3] = static_cast<uint8_t>(adler);
        if (L.raw == 0) {  // the single empty block has no first byte to write its header
            const uint8_t hdr[5] = {1, 0, 0, 0xff, 0xff};
            for (int k = 0; k < 5; ++k) z[2 + k] = hdr[k];
        }
    }
}

// Z^(2^i), filled once per device from shiftTable() (zShiftReady in io.hip). The kernels copy
// it to LDS first: a shift's matrix index depends on each lane's length, and per-lane reads of
// constant memory are global loads whose latency the 32-step product would pay one by one.
static __constant__ uint32_t cZ[kShiftBits][32];

__device__ inline void loadShifts(uint32_t (*sZ)[32]) {
    for (int k = threadIdx.x; k < kShiftBits * 32; k += blockDim.x) (&sZ[0][0])[k] = (&cZ[0][0])[k];
}

// Branch-free: every column is read and masked by its bit of v (a select on the bit became a
// per-lane branch over the wavefront's lanes).
__device__ inline uint32_t gfApplyDev(const uint32_t* z, uint32_t v) {
    uint32_t r = 0;
#pragma unroll
    for (int j = 0; j < 32; ++j) r ^= z[j] & (0u - ((v >> j) & 1u));
    return r;
}

// Z^len v, len < 2^kShiftBits, with the matrices of len's set bits.
__device__ inline uint32_t gfShift(const uint32_t (*sZ)[32], uint32_t v, uint64_t len) {
    for (int i = 0; len; ++i, len >>= 1)
        if (len & 1u) v = gfApplyDev(sZ[i], v);
    return v;
}

// Tree combination of (register, length) pairs over a workgroup of 256 in LDS; after it,
// sReg[0] and sLen[0] are those of the concatenation in thread order.
__device__ inline void combineTree(const uint32_t (*sZ)[32], uint32_t* sReg, uint64_t* sLen) {
    for (int s = 1; s < 256; s <<= 1) {
        __syncthreads();
        const int t = threadIdx.x;
        if ((t % (2 * s)) == 0 && t + s < 256) {
            sReg[t] = gfShift(sZ, sReg[t], sLen[t + s]) ^ sReg[t + s];
            sLen[t] += sLen[t + s];
        }
    }
    __syncthreads();
}

// CRC registers (from zero) of bytes[0, n), one 256-byte piece per thread, combined per
// workgroup into reg[wg], len[wg]. Four bytes per step with four tables (slicing by four:
// T1..T3 are T0 advanced by one to three more zero bytes), so the chain of dependent table
// reads is a quarter of the bytes long. (Staging each wavefront's pieces in LDS with coalesced
// loads first measured slower: 71 against 48 us for a 1080p frame.)
__global__ __launch_bounds__(256) void k_crcPieces(const uint8_t* __restrict__ bytes, uint64_t n,
                                                   uint32_t* __restrict__ reg, uint64_t* __restrict__ len) {
    __shared__ uint32_t sT[4][256];
    __shared__ uint32_t sReg[256];
    __shared__ uint64_t sLen[256];
    __shared__ uint32_t sZ[kShiftBits][32];
    loadShifts(sZ);
    {
        uint32_t c = threadIdx.x;
        for (int k = 0; k < 8; ++k) c = (c & 1u) ? kCrcPoly ^ (c >> 1) : c >> 1;
        sT[0][threadIdx.x] = c;
    }
    __syncthreads();
    for (int k = 1; k < 4; ++k) {
        const uint32_t prev = sT[k - 1][threadIdx.x];
        sT[k][threadIdx.x] = sT[0][prev & 0xffu] ^ (prev >> 8);
        __syncthreads();
    }
    const uint64_t first = (static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x) * kCrcPiece;
    const uint64_t last = min(n, first + kCrcPiece);
    uint32_t r = 0;
    uint64_t i = first;
    for (; i + 4 <= last; i += 4) {
        const uint32_t w = static_cast<uint32_t>(bytes[i]) | (static_cast<uint32_t>(bytes[i + 1]) << 8) |
                           (static_cast<uint32_t>(bytes[i + 2]) << 16) | (static_cast<uint32_t>(bytes[i + 3]) << 24);
        const uint32_t c = r ^ w;
        r = sT[3][c & 0xffu] ^ sT[2][(c >> 8) & 0xffu] ^ sT[1][(c >> 16) & 0xffu] ^ sT[0][c >> 24];
    }
    for (; i < last; ++i) r = sT[0][(r ^ bytes[i]) & 0xffu] ^ (r >> 8);
    sReg[threadIdx.x] = r;
    sLen[threadIdx.x] = last > first ? last - first : 0;
    combineTree(sZ, sReg, sLen);
    if (threadIdx.x == 0) {
        reg[blockIdx.x] = sReg[0];
        len[blockIdx.x] = sLen[0];
    }
}

// One workgroup: combines the workgroup registers in order (256 at a time), applies the initial
// register 0xffffffff over the whole length and the final inversion, and stores the CRC
// big-endian at crcOut.
__global__ __launch_bounds__(256) void k_crcFinish(const uint32_t* __restrict__ reg, const uint64_t* __restrict__ len,
                                                   uint32_t parts, uint8_t* __restrict__ crcOut) {
    __shared__ uint32_t sReg[256];
    __shared__ uint64_t sLen[256];
    __shared__ uint32_t sZ[kShiftBits][32];
    loadShifts(sZ);
    uint32_t acc = 0;
    uint64_t total = 0;
    for (uint32_t base = 0; base < parts; base += 256) {
        const uint32_t i = base + threadIdx.x;
        sReg[threadIdx.x] = i < parts ? reg[i] : 0u;
        sLen[threadIdx.x] = i < parts ? len[i] : 0u;
        combineTree(sZ, sReg, sLen);
        acc = gfShift(sZ, acc, sLen[0]) ^ sReg[0];
        total += sLen[0];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const uint32_t crc = ~(gfShift(sZ, 0xffffffffu, total) ^ acc);
        crcOut[0] = static_cast<uint8_t>(crc >> 24);
        crcOut[1] = static_cast<uint8_t>(crc >> 16);
        crcOut[2] = static_cast<uint8_t>(crc >> 8);
        crcOut[3] = static_cast<uint8_t>(crc);
    }
}

// Workspace of the device encoder for a w x h frame: Adler-32 partial sums, then the CRC
// pieces' registers and lengths.
struct DevWork {
    size_t bodyBlocks, crcBlocks, adlerOff, regOff, lenOff, bytes;
    explicit DevWork(const Layout& L) {
        const size_t perBody = static_cast<size_t>(kBodyBlock) * kBodyItems;
        bodyBlocks = std::max<size_t>(1, (L.raw + perBody - 1) / perBody);
        const size_t perCrc = static_cast<size_t>(256) * kCrcPiece;
        crcBlocks = std::max<size_t>(1, (L.zlib + 4 + perCrc - 1) / perCrc);
        adlerOff = 0;
        regOff = (8 * bodyBlocks + 255) / 256 * 256;
        lenOff = (regOff + 4 * crcBlocks + 255) / 256 * 256;
        bytes = lenOff + 8 * crcBlocks;
    }
};

}  // namespace png
}  // namespace ans
