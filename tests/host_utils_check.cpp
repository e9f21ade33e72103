// host_utils_check.cpp -- test helper for tests/test_host_utils.py: links
// the host utility API of n-body-code_amd/csrc/nbody.hpp (as code written
// against the reference's include/utils.h would) and prints, for each
// whitespace-separated hex bit pattern on stdin (16 digits: a double, 8: a
// float), the bit pattern of InvSqrt2D / InvSqrt2 of it; then two timer()
// readings.
#include <cinttypes>
#include <cstdio>
#include <cstring>
#include <string>
#include <iostream>

#include "nbody.hpp"

int main()
{
    std::string tok;
    while (std::cin >> tok) {
        if (tok.size() == 16) {
            uint64_t b = std::stoull(tok, nullptr, 16);
            double x;
            std::memcpy(&x, &b, 8);
            const double y = InvSqrt2D(x);
            std::memcpy(&b, &y, 8);
            std::printf("%016" PRIx64 "\n", b);
        } else {
            uint32_t b = (uint32_t)std::stoul(tok, nullptr, 16);
            float x;
            std::memcpy(&x, &b, 4);
            const float y = InvSqrt2(x);
            std::memcpy(&b, &y, 4);
            std::printf("%08" PRIx32 "\n", b);
        }
    }
    const double t0 = timer(), t1 = timer();
    std::printf("timer %.6f %.6f\n", t0, t1);
    std::printf("estimate_err %.17g\n", estimate_err(-0.2475, -0.25));
    return 0;
}
