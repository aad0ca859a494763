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

// The polynomial coefficients as one array: A0..A4 of the table path at 0..4,
// B1..B10 of the path near 1 at 5..14 (B0 = -1/2 is exact and stays a
// constant). The kernels read them from a copy in LDS, at the point of use:
// as literals, an FMA with two constant operands needs one of them in a vector
// register, which the compiler would keep for the whole kernel.
constexpr int kLogCoefs = 16;
constexpr double kLogCoef[kLogCoefs] = {
    kLogA[0], kLogA[1], kLogA[2], kLogA[3], kLogA[4], kLogB[1], kLogB[2], kLogB[3],
    kLogB[4], kLogB[5], kLogB[6], kLogB[7], kLogB[8], kLogB[9], kLogB[10], 0.0};
constexpr uint64_t kLogNear1Lo = 0x3fee000000000000ull; // bits of 1 - 2^-4
constexpr uint64_t kLogNear1Hi = 0x3ff1090000000000ull; // bits of 1 + 0x1.09p-4

__host__ __device__ inline double log_near1(double x, const double *c)
{
    const double r = x - 1.0, r2 = r * r, r3 = r * r2;
    const double q3 = __builtin_fma(r3, c[14], __builtin_fma(r2, c[13], __builtin_fma(r, c[12], c[11])));
    const double q2 = __builtin_fma(r3, q3, __builtin_fma(r2, c[10], __builtin_fma(r, c[9], c[8])));
    const double q1 = __builtin_fma(r3, q2, __builtin_fma(r2, c[7], __builtin_fma(r, c[6], c[5])));
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
__host__ __device__ inline double log_table(uint64_t ix, const LogEntry *tab, const double *coef)
{
    const uint64_t tmp = ix - 0x3fe6000000000000ull;
    const int i = (int)((tmp >> (52 - kLogTableBits)) & 127u);
    const int k = (int)((int64_t)tmp >> 52);
    const double z = __builtin_bit_cast(double, ix - (tmp & (0xfffull << 52)));
    const LogEntry e = tab[i];
    const double r = __builtin_fma(z, e.invc, -1.0);
    const double kd = (double)k;
    const double w = __builtin_fma(kd, kLogLn2Hi, e.logc); // exact
    const double hi = w + r;
    const double lo = __builtin_fma(kd, kLogLn2Lo, (w - hi) + r);
    const double r2 = r * r;
    const double p = __builtin_fma(r2, __builtin_fma(r, coef[4], coef[3]),
                                   __builtin_fma(r, coef[2], coef[1]));
    return __builtin_fma(r * r2, p, __builtin_fma(r2, coef[0], lo)) + hi;
}

// log(x) for any x (the decoder forms only positive normal arguments: 1 +
// exp(v) and probability sums floored at EPSILON; zero, subnormals,
// negatives, infinities and NaN follow the library too).
__host__ __device__ inline double log_pos(double x, const LogEntry *tab, const double *coef)
{
    uint64_t ix = __builtin_bit_cast(uint64_t, x);
    if (ix - kLogNear1Lo < kLogNear1Hi - kLogNear1Lo) return log_near1(x, coef);
    const uint32_t top = (uint32_t)(ix >> 48);
    if (top - 0x0010u >= 0x7ff0u - 0x0010u) {
        if ((ix << 1) == 0) return -__builtin_inf();
        if (ix == 0x7ff0000000000000ull) return x;
        if ((top & 0x8000u) || (top & 0x7ff0u) == 0x7ff0u) return (x - x) / (x - x);
        ix = __builtin_bit_cast(uint64_t, x * 0x1p52) - (52ull << 52); // subnormal
    }
    return log_table(ix, tab, coef);
}

// The same with the coefficients from kLogCoef (host callers).
__host__ __device__ inline double log_pos(double x, const LogEntry *tab)
{
    return log_pos(x, tab, kLogCoef);
}

} // namespace sg
