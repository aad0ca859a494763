"""Generate sg_exp_table.h: 2^(j/64) as exact hi + lo double pairs, the ln2/64
split used by the device exp, and the coefficients of its polynomial.

The polynomial: exp(r) - 1 = r + r^2 (c2 + c3 r + c4 r^2 + c5 r^3) for
|r| <= ln2/128, the cubic interpolating (exp(r) - 1 - r)/r^2 at the four
Chebyshev nodes of that interval (near-minimax; truncation below 2^-57.6
relative to exp(r)).

    python sogrand_amd/csrc/make_exp_table.py > sogrand_amd/csrc/sg_exp_table.h
"""
import math
from decimal import Decimal, getcontext

getcontext().prec = 60
N = 64


def cubic_coefficients():
    h = Decimal(2).ln() / (2 * N)
    nodes = [h * Decimal(math.cos((2 * k + 1) * math.pi / 8)) for k in range(4)]

    def power(x, j):
        out = Decimal(1)
        for _ in range(j):
            out *= x
        return out

    m = [[power(x, j) for j in range(4)] + [(x.exp() - 1 - x) / (x * x)] for x in nodes]
    for c in range(4):  # Gauss-Jordan elimination with partial pivoting
        piv = max(range(c, 4), key=lambda r: abs(m[r][c]))
        m[c], m[piv] = m[piv], m[c]
        for r in range(4):
            if r != c:
                f = m[r][c] / m[c][c]
                m[r] = [m[r][k] - f * m[c][k] for k in range(5)]
    return [float(m[i][4] / m[i][i]) for i in range(4)]


def main() -> None:
    rows = []
    for j in range(N):
        v = Decimal(2) ** (Decimal(j) / N)
        hi = float(v)
        lo = float(v - Decimal(hi))
        rows.append(f"    {{{hi.hex()}, {lo.hex()}}}, // 2^({j}/{N})")
    l1 = float(Decimal(2).ln() / N)
    l2 = float(Decimal(2).ln() / N - Decimal(l1))
    inv = float(N / Decimal(2).ln())
    c2, c3, c4, c5 = cubic_coefficients()
    body = "\n".join(rows)
    print(f"""// Table for the device exp: 2^(j/{N}) = hi + lo to about 106 bits, the
// reduction constants and the polynomial coefficients, generated with 60-digit
// decimal arithmetic by make_exp_table.py.
#pragma once

namespace sg {{

constexpr int kExpTableBits = 6;
constexpr double kExpInvLn2N = {inv.hex()}; // {N} / ln 2
constexpr double kExpLn2NHi = {l1.hex()};   // ln 2 / {N}, rounded
constexpr double kExpLn2NLo = {l2.hex()};   // ln 2 / {N} - kExpLn2NHi
// exp(r) - 1 = r + r^2 (c2 + c3 r + c4 r^2 + c5 r^3), |r| <= ln 2 / {2 * N}
constexpr double kExpC2 = {c2.hex()};
constexpr double kExpC3 = {c3.hex()};
constexpr double kExpC4 = {c4.hex()};
constexpr double kExpC5 = {c5.hex()};

struct ExpPair {{
    double hi, lo;
}};

constexpr ExpPair kExp2Table[{N}] = {{
{body}
}};

}} // namespace sg""")


if __name__ == "__main__":
    main()
