// exp(x) of the line decoder (the guess weights exp(-pm), JacLog and the soft
// output) with the arithmetic of the C library the reference calls: the
// table-driven exp of Arm Optimized Routines (N = 128, degree-5 polynomial;
// glibc's dbl-64 exp since 2.28), evaluated with the fused multiply-adds of that
// library's x86-64 FMA build, so every result is the library's bit for bit.
// The same source builds for the device (table in LDS) and for the host, where
// the CPU tests compare it with the C library's exp.
#pragma once

#include <stdint.h>

#include "sg_exp_table.h"

#pragma clang fp contract(off)

namespace sg {

// One 16-byte entry of the LDS exp table. Entries 0..127 hold T[j] and the
// scale bits of 2^(j/128) (as a double's bits); entries 128 and 129 hold the
// polynomial coefficients (C5, C4) and (C3, C2), read at one address by every
// lane (a broadcast), which keeps them out of loop-invariant registers.
struct ExpEntry {
    double a, b;
};
constexpr int kExpTableEntries = 130;

// Entry i of that table.
__host__ __device__ constexpr ExpEntry exp_table_entry(int i)
{
    return i < 128 ? ExpEntry{kExp2Table[i].tail, __builtin_bit_cast(double, kExp2Table[i].sbits)}
           : i == 128 ? ExpEntry{kExpC5, kExpC4}
                      : ExpEntry{kExpC3, kExpC2};
}

// The reduction x = k ln2/128 + r (k from one FMA with the 1.5 2^52 shifter,
// whose low word is then k), the table lookup and the polynomial:
// exp(x) = scale (1 + tmp) with scale = 2^(k/128) as bits.
struct ExpParts {
    double tmp;
    uint64_t sbits; // bits of 2^(k/128) (exponent may be out of range for |x| >= 512)
    uint64_t ki;    // low word: k
};

// Scale bits T[j] bits + (k << 45). The addend's low word is zero, so only the
// high words add: on the device one 32-bit shift-add written as the
// instruction (as C, the compiler rebuilds a 64-bit add with register moves).
__host__ __device__ inline uint64_t scale_bits(double tb, uint64_t ki)
{
#if defined(__HIP_DEVICE_COMPILE__)
    struct Words {
        uint32_t lo, hi;
    };
    Words w = __builtin_bit_cast(Words, tb);
    asm("v_lshl_add_u32 %0, %1, 13, %0" : "+v"(w.hi) : "v"((uint32_t)ki));
    return __builtin_bit_cast(uint64_t, w);
#else
    return __builtin_bit_cast(uint64_t, tb) + (ki << 45);
#endif
}

__host__ __device__ inline ExpParts exp_parts(double x, const ExpEntry *tab)
{
    const double kd0 = __builtin_fma(x, kExpInvLn2N, kExpShift);
    const uint64_t ki = __builtin_bit_cast(uint64_t, kd0);
    const double kd = kd0 - kExpShift;
    const double r = __builtin_fma(kd, kExpNegLn2LoN, __builtin_fma(kd, kExpNegLn2HiN, x));
    const ExpEntry t = tab[ki & 127u];
    const ExpEntry c54 = tab[128], c32 = tab[129];
    const double r2 = r * r;
    const double tmp = __builtin_fma(r2 * r2, __builtin_fma(r, c54.a, c54.b),
                                     __builtin_fma(r2, __builtin_fma(r, c32.a, c32.b), t.a + r));
    return {tmp, scale_bits(t.b, ki), ki};
}

// |x| < 512: scale + scale tmp, one fused multiply-add.
__host__ __device__ inline double exp_main(double x, const ExpEntry *tab)
{
    const ExpParts p = exp_parts(x, tab);
    const double scale = __builtin_bit_cast(double, p.sbits);
    return __builtin_fma(scale, p.tmp, scale);
}

// |x| >= 512, infinities and NaN. For k > 0 the scale is formed 2^1009 lower
// and multiplied back; for k < 0 it is formed 2^1022 higher, the sum is formed
// unfused (the library's special case sits in its own block, where its
// compiler does not contract), and a result below 2^-1022 is rounded once, at
// its subnormal precision, before it is scaled down.
__host__ __device__ inline double exp_large(double x, const ExpEntry *tab)
{
    if (!(__builtin_fabs(x) < 1024.0)) {
        if (x != x) return 1.0 + x;
        return x > 0.0 ? __builtin_inf() : 0.0;
    }
    const ExpParts p = exp_parts(x, tab);
    if ((p.ki & 0x80000000u) == 0) {
        const double scale = __builtin_bit_cast(double, p.sbits - (1009ull << 52));
        return 0x1p1009 * __builtin_fma(scale, p.tmp, scale);
    }
    const double scale = __builtin_bit_cast(double, p.sbits + (1022ull << 52));
    const double st = scale * p.tmp;
    double y = scale + st;
    if (y < 1.0) {
        double lo = (scale - y) + st;
        const double hi = 1.0 + y;
        lo = ((1.0 - hi) + y) + lo;
        y = (hi + lo) - 1.0;
    }
    return 0x1p-1022 * y;
}

// exp(x) for any x.
__host__ __device__ inline double exp_of(double x, const ExpEntry *tab)
{
    double y = exp_main(x, tab);
    if (!(__builtin_fabs(x) < 512.0)) y = exp_large(x, tab);
    return y;
}

} // namespace sg
