// The stopping rule's confidence test at a list insertion (sogrand_main_logic
// :634, :676, :716 with getLConf :767-774):
//
//     conf = (P_pos + P_neg > EPSILON) ? P_pos / (P_pos + P_neg) : 1.0
//     stop if conf > thres
//
// decided without the division in all but rare cases, with the same outcome
// for every input. Shared by the host (the CPU test compares it with the
// division on adversarial inputs) and the device (list insertion).
#pragma once

#include <stdint.h>

#include "sg_common.h"

#pragma clang fp contract(off)

namespace sg {

// Half the gap between thres and the next larger double, for conf_above; NaN
// (conf_above then divides every time) where the argument below does not hold:
// thres outside [2^-60, 2^1000) or eps below 2^-900. For thres = 1 - 1e-5
// it is 2^-54.
SG_HD inline double conf_gap(double thres, double eps)
{
    if (!(thres >= 0x1p-60 && thres < 0x1p1000 && eps >= 0x1p-900)) return __builtin_nan("");
    const double next = __builtin_bit_cast(double, __builtin_bit_cast(uint64_t, thres) + 1u);
    return (next - thres) * 0.5;
}

// Whether getLConf's conf exceeds thres. With s = P_pos + P_neg > eps (so
// s > 2^-900 > 0 where gap is a number), the quotient rounds to a double above
// thres iff P_pos / s lies above the midpoint m = thres + gap of thres and its
// successor, or on it with ties-to-even choosing the successor; i.e. iff
// P_pos - thres s > gap s. The fused multiply-add gives the left side rounded
// once, and rounding is monotonic while gap s is an exact double (gap is a
// power of two and s is far above the subnormal range): a result strictly
// above (below) gap s proves the exact difference above (below) it. Equality,
// NaN, infinities and a NaN gap take the division, as the reference does.
SG_HD inline bool conf_above(double ppos, double pneg, double eps, double thres, double gap)
{
    const double s = ppos + pneg;
    const double a = __builtin_fma(-thres, s, ppos);
    const double gs = gap * s;
    const bool cut = !(s > eps);
    bool above = cut ? 1.0 > thres : a > gs;
    // the one branch: rare, so a wave of lanes mostly skips it
    if (!cut && !(a > gs) && !(a < gs)) above = ppos / s > thres;
    return above;
}

} // namespace sg
