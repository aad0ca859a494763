"""Generate sg_log_table.h: the table and constants of the device logarithm
(table-driven, after P. T. P. Tang, "Table-driven implementation of the
logarithm function in IEEE floating-point arithmetic", ACM TOMS 16(4), 1990).

For c = 1 + i/128 (i = -37 .. 53, covering mantissas in [sqrt(1/2), sqrt(2))):
invc = 1/c rounded to double, and -log(invc) split as logc_hi (a multiple of
2^-42, so that k ln2_hi + logc_hi is exact for |k| <= 1075) plus logc_lo.

    python sogrand_amd/csrc/make_log_table.py > sogrand_amd/csrc/sg_log_table.h
"""
from decimal import Decimal, getcontext

getcontext().prec = 80
LO, HI = -37, 53
Q = Decimal(2) ** -42


def split(v: Decimal):
    hi = (v / Q).to_integral_value() * Q
    return float(hi), float(v - hi)


def main() -> None:
    rows = []
    for i in range(LO, HI + 1):
        c = 1 + Decimal(i) / 128
        invc = float(1 / c)
        logc = -Decimal(invc).ln()
        hi, lo = split(logc)
        assert Decimal(hi) == (Decimal(hi) / Q).to_integral_value() * Q
        rows.append(f"    {{{invc.hex()}, {hi.hex()}, {lo.hex()}}}, // c = 1 + {i}/128")
    l2hi, l2lo = split(Decimal(2).ln())
    body = "\n".join(rows)
    print(f"""// Table of the device logarithm: for c = 1 + i/128, i = {LO} .. {HI}, invc = 1/c
// rounded and -log(invc) = logc_hi + logc_lo (logc_hi a multiple of 2^-42),
// generated with 80-digit decimal arithmetic by make_log_table.py.
#pragma once

namespace sg {{

constexpr int kLogTableFirst = {LO};
constexpr int kLogTableEntries = {HI - LO + 1};
constexpr double kLogLn2Hi = {l2hi.hex()}; // ln 2 rounded to a multiple of 2^-42
constexpr double kLogLn2Lo = {l2lo.hex()}; // ln 2 - kLogLn2Hi

struct LogEntry {{
    double invc, logc_hi, logc_lo;
}};

constexpr LogEntry kLogTable[{HI - LO + 1}] = {{
{body}
}};

}} // namespace sg""")


if __name__ == "__main__":
    main()
