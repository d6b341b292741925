// Host check of the multiply-shift division used by the kernels' index math
// (mpsq_device.h: make_fdiv / fdiv): exhaustive over x < 2^16 for every
// divisor d < 2^12, and at the extremes x = q*d - 1, q*d, q*d + d - 1 and the
// largest x < 2^31 for divisors up to 2^31 - 1.  Prints "ok" or the first
// mismatch.
#include <cstdio>
#include <cstdint>

#include "mpsq_device.h"

using mpsq::FDiv;
using mpsq::fdiv;
using mpsq::make_fdiv;

static int check(uint32_t x, uint32_t d, FDiv f)
{
    if (fdiv(x, f) != x / d) {
        std::printf("mismatch: %u / %u -> %u (want %u)\n", x, d, fdiv(x, f), x / d);
        return 1;
    }
    return 0;
}

int main()
{
    for (uint32_t d = 1; d < 4096; ++d) {
        const FDiv f = make_fdiv(d);
        for (uint32_t x = 0; x < 65536; ++x)
            if (check(x, d, f)) return 1;
    }
    const uint32_t xmax = 0x7fffffffu;
    uint64_t s = 88172645463325252ull;
    for (int it = 0; it < 200000; ++it) {
        s ^= s << 13; s ^= s >> 7; s ^= s << 17;
        uint32_t d = (uint32_t)(s % xmax) + 1;
        if (it % 3 == 0) d = (d % 100000) + 1;
        const FDiv f = make_fdiv(d);
        const uint64_t q = (s >> 33) % (xmax / d + 1);
        const uint64_t xs[5] = {q * d - 1, q * d, q * d + d - 1, xmax, (s >> 1) & xmax};
        for (uint64_t x : xs)
            if (x <= xmax && check((uint32_t)x, d, f)) return 1;
    }
    std::printf("ok\n");
    return 0;
}
