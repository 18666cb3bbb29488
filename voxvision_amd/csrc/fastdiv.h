// Division of 31-bit unsigned integers by a divisor fixed per launch, as
// one 32x32->64 multiply and a shift (Granlund and Montgomery, "Division
// by invariant integers using multiplication", PLDI 1994, theorem 4.2 with
// N = 31): for 1 <= d < 2^31, l = ceil(log2 d) and m = ceil(2^(31+l) / d),
// floor(n / d) = floor(n * m / 2^(31+l)) for every 0 <= n < 2^31, and m
// fits in 32 bits.  The GPU has no integer divide instruction; this
// replaces the tens of instructions of a generic 32-bit division.
#pragma once
#include <stdint.h>

struct FastDiv {
    uint32_t d;  // the divisor
    uint32_t m;  // ceil(2^s / d)
    uint32_t s;  // 31 + ceil(log2 d)

    // n / d for n < 2^31
    __host__ __device__ inline uint32_t div(uint32_t n) const {
        return (uint32_t)(((uint64_t)n * m) >> s);
    }
};

// d in [1, 2^31)
static inline FastDiv make_fastdiv(uint32_t d) {
    uint32_t l = 0;
    while ((1ull << l) < d) ++l;
    FastDiv f;
    f.d = d;
    f.s = 31u + l;
    f.m = (uint32_t)(((1ull << f.s) + d - 1ull) / d);
    return f;
}
