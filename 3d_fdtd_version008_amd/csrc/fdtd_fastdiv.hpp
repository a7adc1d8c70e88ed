// fdtd_fastdiv.hpp - division by a divisor fixed before a launch.
//
// A kernel that splits its work-group id into tile coordinates divides by
// the number of tiles along an axis; a run-time 32-bit division costs some
// 25 scalar instructions per quotient on gfx950.  With the divisor known
// when the kernel is launched, the host computes a multiplier once and the
// quotient becomes one 32x32->64 bit multiply and a shift (the method of
// Granlund and Montgomery, "Division by invariant integers using
// multiplication", PLDI 1994, in its round-up form):
//
//   l = ceil(log2 d),  m = ceil(2^(31+l) / d)  (< 2^32 since d > 2^(l-1)),
//   n / d = floor(n * m / 2^(31+l))            for 0 <= n < 2^31.
//
// Proof: m = (2^(31+l) + e) / d with 0 <= e < d <= 2^l, so
// n*m / 2^(31+l) = n/d + n*e / (d * 2^(31+l)), and n*e < 2^31 * 2^l makes
// the second term smaller than 1/d; with n = q*d + r, r <= d-1, the sum
// stays below q + 1.
//
// The functions are constexpr (and so usable in device code as well) and
// need no HIP header, so the host test compiles this file with g++ alone.
#pragma once
#include <stdint.h>

namespace fdtdk {

struct FastDiv {
    unsigned d;   // divisor (>= 1)
    unsigned m;   // multiplier
    unsigned sh;  // 31 + ceil(log2 d)
};

constexpr FastDiv make_fastdiv(unsigned d) {
    unsigned l = 0;
    while (l < 31 && (1u << l) < d) ++l;
    const uint64_t m = ((uint64_t(1) << (31 + l)) + d - 1) / d;
    return FastDiv{d, (unsigned)m, 31 + l};
}

// n / f.d for 0 <= n < 2^31
constexpr unsigned fdiv(unsigned n, const FastDiv& f) {
    return (unsigned)(((uint64_t)n * f.m) >> f.sh);
}

}  // namespace fdtdk
