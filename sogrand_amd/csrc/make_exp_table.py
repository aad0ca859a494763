"""Generate sg_exp_table.h: the table and constants of the decoder's exp.

The method is the table-driven exp of Arm Optimized Routines (Szabolcs Nagy,
2018; MIT / Apache-2.0), which is also the exp of the C library the reference
calls (glibc 2.28 and later, sysdeps/ieee754/dbl-64/e_exp.c). With N = 128:

    x = k ln2/N + r, |r| <= ln2/(2N),   k = round(x N/ln2)
    2^(k/N) = 2^(k>>7) * H[j] * (1 + T[j]),   j = k mod 128
    exp(x) = scale + scale * (T[j] + r + r^2 (C2 + r C3) + r^4 (C4 + r C5))

The table holds, per j, the tail T[j] = ((2^(j/N) - H[j]) / H[j]) rounded, with
H[j] = 2^(j/N) rounded to double, and the bits of H[j] minus j << 45, so that
adding k << 45 to them gives the bits of 2^(k/N) rounded. Both are computed
here with 80-digit decimal arithmetic. The reduction constants and the
polynomial coefficients C2..C5 are the published ones of that design.

    python sogrand_amd/csrc/make_exp_table.py > sogrand_amd/csrc/sg_exp_table.h
"""
import struct
from decimal import Decimal, getcontext

getcontext().prec = 80
N = 128


def bits(x: float) -> int:
    return struct.unpack("<Q", struct.pack("<d", x))[0]


def main() -> None:
    rows = []
    for j in range(N):
        v = Decimal(2) ** (Decimal(j) / N)
        h = float(v)
        tail = float((v - Decimal(h)) / Decimal(h))
        sbits = (bits(h) - (j << 45)) & 0xFFFFFFFFFFFFFFFF
        rows.append(f"    {{{tail.hex()}, 0x{sbits:016x}ull}}, // 2^({j}/{N})")
    body = "\n".join(rows)
    print(f"""// Table of the decoder's exp (the table-driven exp of Arm Optimized Routines,
// N = {N}): per j, the tail T[j] of 2^(j/{N}) = H[j] (1 + T[j]) and the bits of
// H[j] minus j << 45, generated with 80-digit decimal arithmetic by
// make_exp_table.py; the reduction constants and polynomial coefficients of
// that design.
#pragma once

#include <stdint.h>

namespace sg {{

constexpr int kExpTableBits = 7;
constexpr double kExpInvLn2N = 0x1.71547652b82fep0 * {N}; // {N} / ln 2
constexpr double kExpNegLn2HiN = -0x1.62e42fefa0000p-8;  // -ln 2 / {N}, high part
constexpr double kExpNegLn2LoN = -0x1.cf79abc9e3b3ap-47; // -ln 2 / {N}, low part
constexpr double kExpShift = 0x1.8p52;                   // round-to-integer shifter
// exp(r) - 1 - T ~ r + r^2 (C2 + r C3) + r^4 (C4 + r C5), |r| <= ln 2 / {2 * N}
constexpr double kExpC2 = 0x1.ffffffffffdbdp-2;
constexpr double kExpC3 = 0x1.555555555543cp-3;
constexpr double kExpC4 = 0x1.55555cf172b91p-5;
constexpr double kExpC5 = 0x1.1111167a4d017p-7;

struct ExpTail {{
    double tail;
    uint64_t sbits;
}};

constexpr ExpTail kExp2Table[{N}] = {{
{body}
}};

}} // namespace sg""")


if __name__ == "__main__":
    main()
