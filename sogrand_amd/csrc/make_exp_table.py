"""Generate sg_exp_table.h: 2^(j/32) as exact hi + lo double pairs and the
ln2/32 split used by the device exp of the guessing loop.

    python sogrand_amd/csrc/make_exp_table.py > sogrand_amd/csrc/sg_exp_table.h
"""
from decimal import Decimal, getcontext

getcontext().prec = 60
N = 32


def main() -> None:
    rows = []
    for j in range(N):
        v = Decimal(2) ** (Decimal(j) / N)
        hi = float(v)
        lo = float(v - Decimal(hi))
        rows.append(f"    {{{hi.hex()}, {lo.hex()}}}, // 2^({j}/32)")
    l1 = float(Decimal(2).ln() / N)
    l2 = float(Decimal(2).ln() / N - Decimal(l1))
    inv = float(N / Decimal(2).ln())
    body = "\n".join(rows)
    print(f"""// Table for the device exp of the guessing loop: 2^(j/32) = hi + lo to about
// 106 bits, generated with 60-digit decimal arithmetic by make_exp_table.py.
#pragma once

namespace sg {{

constexpr double kExpInvLn2N = {inv.hex()}; // 32 / ln 2
constexpr double kExpLn2NHi = {l1.hex()};   // ln 2 / 32, rounded
constexpr double kExpLn2NLo = {l2.hex()};   // ln 2 / 32 - kExpLn2NHi

struct ExpPair {{
    double hi, lo;
}};

constexpr ExpPair kExp2Table[32] = {{
{body}
}};

}} // namespace sg""")


if __name__ == "__main__":
    main()
