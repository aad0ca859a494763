// exp(x) of the line decoder (the guess weights exp(-pm), JacLog and the soft
// output), table-driven after P. T. P. Tang (ACM TOMS 15(2), 1989). The same
// source builds for the device (table in LDS) and for the host, where the CPU
// tests measure it against the C library's exp.
#pragma once

#include <stdint.h>

#include "sg_exp_table.h"

#pragma clang fp contract(off)

namespace sg {

// One 16-byte entry of the LDS exp table. Entries 0..63 hold 2^(j/64) = a + b
// (hi, lo); entries 64 and 65 hold the polynomial coefficients (c5, c4) and
// (c3, c2), read at one address by every lane (a broadcast). Reading the
// coefficients from LDS keeps them out of loop-invariant registers, which the
// compiler otherwise pays for with a copy per step; as a separate entry rather
// than a copy in every entry, the per-lane read at a data-dependent index
// stays 16 bytes wide.
struct ExpEntry {
    double a, b;
};
constexpr int kExpTableEntries = 66;

// Entry i of that table.
__host__ __device__ constexpr ExpEntry exp_table_entry(int i)
{
    return i < 64 ? ExpEntry{kExp2Table[i].hi, kExp2Table[i].lo}
           : i == 64 ? ExpEntry{kExpC5, kExpC4}
                     : ExpEntry{kExpC3, kExpC2};
}

// exp(x) for x in [-746, 710], table-driven after Tang: x = (64k + j) ln2/64 + r
// with |r| <= ln2/128, exp(r) - 1 = q = r + r^2 (c2 + c3 r + c4 r^2 + c5 r^3)
// (near-minimax, see make_exp_table.py), and 2^(j/64) = hi + lo from the LDS
// table; the result hi + (hi q + lo (1 + q)) scaled by 2^k. The nearest
// integer to x 64/ln2 comes from one FMA with 1.5 2^52, whose low word is then
// that integer. Within 0.54 ulp; it differs from the correctly rounded value
// on about 0.6 % of inputs. Results below 2^-1075 (x below about -745.13)
// round to 0 in the final ldexp, as exp's do.
__host__ __device__ inline double exp_of(double x, const ExpEntry *tab)
{
    const double shifter = 0x1.8p52;
    const double tk = __builtin_fma(x, kExpInvLn2N, shifter);
    const double kd = tk - shifter;
    const int k = (int)(uint32_t)__builtin_bit_cast(uint64_t, tk);
    const ExpEntry t = tab[k & 63];
    const ExpEntry c54 = tab[64], c32 = tab[65];
    double r = __builtin_fma(-kd, kExpLn2NHi, x);
    r = __builtin_fma(-kd, kExpLn2NLo, r);
    double q = c54.a;
    q = __builtin_fma(q, r, c54.b);
    q = __builtin_fma(q, r, c32.a);
    q = __builtin_fma(q, r, c32.b);
    q = __builtin_fma(q, r, 1.0);
    q = q * r;
    const double s = __builtin_fma(t.b, q, t.b);
    const double m = __builtin_fma(t.a, q, s) + t.a;
    return __builtin_ldexp(m, k >> 6);
}

} // namespace sg
