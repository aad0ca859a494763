// Natural logarithm for the line decoder's set-up (JacLog) and soft output,
// table-driven after P. T. P. Tang (ACM TOMS 16(4), 1990). The same source
// builds for the device (table in LDS) and for the host (the constant table),
// where the CPU tests measure it against the C library's log.
//
// x = 2^k m with m in [sqrt(1/2), sqrt(2)); c = 1 + i/128 is the table point
// nearest m, invc = 1/c rounded and logc = -log(invc) (so log(m) =
// logc + log(m invc) exactly). m invc = p_hi + p_lo exactly (a product and
// its FMA residual) and r = p_hi - 1 is exact (Sterbenz), |r| < 0.0056, so
//   log(x) = k ln2 + logc + r + p_lo (1 - r) + (log1p(r) - r)
// with log1p(r) - r by its Taylor series to r^8 (truncation below 2^-62 of
// the result). k ln2_hi + logc_hi is exact (both multiples of 2^-42), its
// sum with r is split into s + err exactly (Fast2Sum: |hi| >= |r| or hi = 0),
// and the small terms are added before the one final rounding s + lo.
// Defined for positive normal finite x, the only arguments the decoder forms
// (1 + exp(v) with |v| <= 50, and probability sums floored at 1e-40 or
// more); the range is not checked, which keeps the code small.
#pragma once

#include <math.h>

#include "sg_log_table.h"

#pragma clang fp contract(off)

namespace sg {

__host__ __device__ inline double log_pos(double x, const LogEntry *tab)
{
    int e;
    const double mant = frexp(x, &e); // [0.5, 1)
    const bool low = mant < 0x1.6a09e667f3bcdp-1; // sqrt(1/2), rounded up
    const double m = low ? 2.0 * mant : mant;   // [sqrt(1/2), sqrt(2)), exact
    const int k = low ? e - 1 : e;
    const int i = (int)rint(fma(m, 128.0, -128.0)); // (m - 1) 128 is exact
    const LogEntry c = tab[i - kLogTableFirst];
    const double ph = m * c.invc;
    const double pl = fma(m, c.invc, -ph);
    const double r = ph - 1.0;
    const double r2 = r * r;
    double q = -0x1p-3;                       // -1/8
    q = fma(q, r, 0x1.2492492492492p-3);      // 1/7
    q = fma(q, r, -0x1.5555555555555p-3);     // -1/6
    q = fma(q, r, 0x1.999999999999ap-3);      // 1/5
    q = fma(q, r, -0x1p-2);                   // -1/4
    q = fma(q, r, 0x1.5555555555555p-2);      // 1/3
    q = fma(q, r, -0x1p-1);                   // -1/2
    q = q * r2;                               // log1p(r) - r
    const double kd = (double)k;
    const double hi = fma(kd, kLogLn2Hi, c.logc_hi); // exact
    const double s = hi + r;
    const double err = (hi - s) + r;
    double lo = fma(kd, kLogLn2Lo, c.logc_lo) + err;
    lo = lo + fma(-r, pl, pl);
    lo = lo + q;
    return s + lo;
}

} // namespace sg
