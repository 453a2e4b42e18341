// TEST INFRASTRUCTURE ONLY. Checks the device-side expf restatement (aniso_amd/csrc/activate.hpp,
// run here on the host) against the C library's expf, which is what the reference PLY loader
// calls (std::exp(float), src/ply_loader.cpp:124-125): every `step`-th float32 bit pattern is
// compared, NaNs by class. Prints the number compared and the number that differ.
#include "activate.hpp"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

int main(int argc, char** argv) {
    const uint64_t step = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 7;
    uint64_t compared = 0, differ = 0;
    for (uint64_t b = 0; b < (uint64_t(1) << 32); b += step) {
        const uint32_t bits = static_cast<uint32_t>(b);
        float x;
        std::memcpy(&x, &bits, 4);
        const float want = std::exp(x), got = ans::expfTable(x);
        uint32_t wb, gb;
        std::memcpy(&wb, &want, 4);
        std::memcpy(&gb, &got, 4);
        ++compared;
        if (wb != gb && !(std::isnan(want) && std::isnan(got))) {
            if (differ < 5) std::printf("x=%a libm=%a restated=%a\n", x, want, got);
            ++differ;
        }
    }
    std::printf("compared %llu differ %llu\n", static_cast<unsigned long long>(compared),
                static_cast<unsigned long long>(differ));
    return differ == 0 ? 0 : 1;
}
