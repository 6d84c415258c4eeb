// fastdiv.h -- division of 31-bit unsigned integers by a run-time invariant
// divisor with one high multiply, an add and a shift (the "round-up" method of
// T. Granlund and P. Montgomery, "Division by invariant integers using
// multiplication", PLDI 1994, section 4).
//
// For a divisor d >= 1 let s = ceil(log2 d) and m' = floor(2^(32+s) / d) + 1.
// Then for every n < 2^31, floor(n / d) = floor(n * m' / 2^(32+s)): writing
// m' = (2^(32+s) + e) / d with 0 < e <= d, the quotient's error term is
// n * e / (d * 2^(32+s)) < 2^-(s+1) < 1 / d, too small to cross an integer.
// m' lies strictly between 2^32 and 2^33 (s <= 31), so m = m' - 2^32 fits 32
// bits, and n * m' / 2^32 = hi(n * m) + n stays below 2^32 because n < 2^31.
//
// Plain C++ (no HIP headers) so that host code and a host-only test can use
// it; in device code the 64-bit product's high half becomes v_mul_hi_u32.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define SIPGPU_HD __host__ __device__
#else
#define SIPGPU_HD
#endif

namespace sipgpu {

struct FastDiv {
    uint32_t d; // the divisor (>= 1)
    uint32_t m; // m' - 2^32
    uint32_t s; // ceil(log2 d)
};

inline FastDiv make_fastdiv(uint32_t d) {
    uint32_t s = 0;
    while ((uint64_t(1) << s) < d) ++s;
    const uint64_t m = ((uint64_t(1) << 32) * ((uint64_t(1) << s) - d)) / d + 1;
    return FastDiv{d, uint32_t(m), s};
}

// floor(n / f.d) for n < 2^31.
SIPGPU_HD inline uint32_t fastdiv(uint32_t n, const FastDiv &f) {
    const uint32_t hi = uint32_t((uint64_t(n) * f.m) >> 32);
    return (hi + n) >> f.s;
}

} // namespace sipgpu
