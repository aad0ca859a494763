// Natural logarithm of the line decoder's set-up (JacLog) and soft output, with
// the arithmetic of the C library the reference calls: the table-driven log
// of Arm Optimized Routines (N = 128; glibc's dbl-64 log since 2.28), evaluated
// with the fused multiply-adds of that library's x86-64 FMA build, so every
// result is the library's bit for bit. The same source builds for the device
// (table in LDS) and for the host, where the CPU tests compare it with the C
// library's log.
//
// Table path: x = 2^k z with z in [0x1.6p-1, 0x1.6p0) (exact), subinterval i
// from the top 7 bits of z's offset, r = z invc - 1 by one FMA, and
//   log(x) = (k ln2_hi + logc) + r + k ln2_lo + r^2 A0 + r^3 (A1 + ... + r^3 A4)
// with the first sum split exactly into hi + lo. Near 1 (x - 1 in
// [-2^-4, 0x1.09p-4)) a polynomial in r = x - 1 with the r^2/2 term split so
// that it is exact.
#pragma once

#include <stdint.h>

#include "sg_log_table.h"

#pragma clang fp contract(off)

namespace sg {

constexpr int kLogTableEntries = 128;
constexpr uint64_t kLogNear1Lo = 0x3fee000000000000ull; // bits of 1 - 2^-4
constexpr uint64_t kLogNear1Hi = 0x3ff1090000000000ull; // bits of 1 + 0x1.09p-4

__host__ __device__ inline double log_near1(double x)
{
    const double r = x - 1.0, r2 = r * r, r3 = r * r2;
    const double q3 = __builtin_fma(r3, kLogB[10], __builtin_fma(r2, kLogB[9], __builtin_fma(r, kLogB[8], kLogB[7])));
    const double q2 = __builtin_fma(r3, q3, __builtin_fma(r2, kLogB[6], __builtin_fma(r, kLogB[5], kLogB[4])));
    const double q1 = __builtin_fma(r3, q2, __builtin_fma(r2, kLogB[3], __builtin_fma(r, kLogB[2], kLogB[1])));
    // r = rhi + rlo with rhi of 26 bits, so rhi^2 B0 is exact
    const double w27 = r * 0x1p27;
    const double rhi = (r + w27) - w27, rlo = r - rhi;
    const double w = rhi * rhi * kLogB[0];
    const double hi = r + w;
    double lo = (r - hi) + w;
    lo = __builtin_fma(kLogB[0] * rlo, rhi + r, lo);
    const double y = __builtin_fma(r3, q1, lo);
    return y + hi;
}

// Table path for the bits ix of a positive normal x.
__host__ __device__ inline double log_table(uint64_t ix, const LogEntry *tab)
{
    const uint64_t tmp = ix - 0x3fe6000000000000ull;
    const int i = (int)((tmp >> (52 - kLogTableBits)) & 127u);
    const int k = (int)((int64_t)tmp >> 52);
    const double z = __builtin_bit_cast(double, ix - (tmp & (0xfffull << 52)));
    const LogEntry c = tab[i];
    const double r = __builtin_fma(z, c.invc, -1.0);
    const double kd = (double)k;
    const double w = __builtin_fma(kd, kLogLn2Hi, c.logc); // exact
    const double hi = w + r;
    const double lo = __builtin_fma(kd, kLogLn2Lo, (w - hi) + r);
    const double r2 = r * r;
    const double p = __builtin_fma(r2, __builtin_fma(r, kLogA[4], kLogA[3]),
                                   __builtin_fma(r, kLogA[2], kLogA[1]));
    return __builtin_fma(r * r2, p, __builtin_fma(r2, kLogA[0], lo)) + hi;
}

// log(x) for any x (the decoder forms only positive normal arguments: 1 +
// exp(v) and probability sums floored at EPSILON; zero, subnormals,
// negatives, infinities and NaN follow the library too).
__host__ __device__ inline double log_pos(double x, const LogEntry *tab)
{
    uint64_t ix = __builtin_bit_cast(uint64_t, x);
    if (ix - kLogNear1Lo < kLogNear1Hi - kLogNear1Lo) return log_near1(x);
    const uint32_t top = (uint32_t)(ix >> 48);
    if (top - 0x0010u >= 0x7ff0u - 0x0010u) {
        if ((ix << 1) == 0) return -__builtin_inf();
        if (ix == 0x7ff0000000000000ull) return x;
        if ((top & 0x8000u) || (top & 0x7ff0u) == 0x7ff0u) return (x - x) / (x - x);
        ix = __builtin_bit_cast(uint64_t, x * 0x1p52) - (52ull << 52); // subnormal
    }
    return log_table(ix, tab);
}

} // namespace sg
