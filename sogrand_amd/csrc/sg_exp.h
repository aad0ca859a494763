// exp(x) of the line decoder (the guess weights exp(-pm), JacLog and the soft
// output), table-driven after P. T. P. Tang (ACM TOMS 15(2), 1989). The same
// source builds for the device (table in LDS) and for the host, where the CPU
// tests measure it against the C library's exp.
#pragma once

#include "sg_exp_table.h"

#pragma clang fp contract(off)

namespace sg {

// One 16-byte entry of the LDS exp table. Entries 0..31 hold 2^(j/32) = a + b
// (hi, lo); entries 32 and 33 hold the Taylor coefficients (1/6!, 1/5!) and
// (1/4!, 1/3!) of exp(r) - 1, read at one address by every lane (a broadcast).
// Reading the coefficients from LDS keeps them out of loop-invariant
// registers, which the compiler otherwise pays for with a copy per step; as a
// separate entry rather than a copy in every entry, the per-lane read at a
// data-dependent index stays 16 bytes wide.
struct ExpEntry {
    double a, b;
};
constexpr int kExpTableEntries = 34;

// Entry i of that table.
__host__ __device__ constexpr ExpEntry exp_table_entry(int i)
{
    return i < 32 ? ExpEntry{kExp2Table[i].hi, kExp2Table[i].lo}
           : i == 32 ? ExpEntry{1.0 / 720.0, 1.0 / 120.0}
                     : ExpEntry{1.0 / 24.0, 1.0 / 6.0};
}

// exp(x) for x in [-746, 710], Tang-style: x = (32k + j) ln2/32 + r with
// |r| <= ln2/64, exp(r) - 1 = q by its Taylor polynomial to r^6, and
// 2^(j/32) = hi + lo from the LDS table; the result hi + (hi q + lo (1 + q))
// scaled by 2^k. Within 0.55 ulp; it differs from the correctly rounded value
// on about 1 input in 200. Results below 2^-1075 (x below about -745.13)
// round to 0 in the final ldexp, as exp's do.
__host__ __device__ inline double exp_of(double x, const ExpEntry *tab)
{
    const double kd = __builtin_rint(x * kExpInvLn2N);
    const int k = (int)kd;
    const ExpEntry t = tab[k & 31];
    const ExpEntry c65 = tab[32], c43 = tab[33];
    double r = __builtin_fma(-kd, kExpLn2NHi, x);
    r = __builtin_fma(-kd, kExpLn2NLo, r);
    double q = c65.a;
    q = __builtin_fma(q, r, c65.b);
    q = __builtin_fma(q, r, c43.a);
    q = __builtin_fma(q, r, c43.b);
    q = __builtin_fma(q, r, 0.5);
    q = __builtin_fma(q, r, 1.0);
    q = q * r;
    const double s = __builtin_fma(t.b, q, t.b);
    const double m = __builtin_fma(t.a, q, s) + t.a;
    return __builtin_ldexp(m, k >> 5);
}

} // namespace sg
