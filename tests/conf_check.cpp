// The division-free confidence test (sogrand_amd/csrc/sg_conf.h), built for the
// host from the same source, against getLConf's own expression
// (SOGRAND_C/cubic_decoder_final.c:767-774: the division, then conf > thres) on
// inputs aimed at the decision boundary: quotients within a few ulps of thres
// and of the midpoint above it, exact ties, the epsilon cut, negative and
// non-finite found masses, and several thresholds and epsilons (including ones
// for which the test falls back to the division). Prints
//   conf n mismatches divisions
// Usage: conf_check [n] [seed]
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "sg_conf.h"

static bool reference(double ppos, double pneg, double eps, double thres)
{
    const double conf = (ppos + pneg > eps) ? (ppos / (ppos + pneg)) : 1.0;
    return conf > thres;
}

static bool divides(double ppos, double pneg, double eps, double thres, double gap)
{
    const double s = ppos + pneg;
    if (!(s > eps)) return false;
    const double a = std::fma(-thres, s, ppos), gs = gap * s;
    return !(a > gs) && !(a < gs);
}

int main(int argc, char **argv)
{
    const long n = argc > 1 ? atol(argv[1]) : 2000000;
    uint64_t st = argc > 2 ? strtoull(argv[2], nullptr, 10) : 0x9e3779b97f4a7c15ull;
    auto next = [&]() {
        st ^= st << 13;
        st ^= st >> 7;
        st ^= st << 17;
        return st;
    };
    auto u01 = [&]() { return (double)(next() >> 11) * 0x1p-53; };
    const double thresholds[] = {1.0 - 1e-5, 1.0 - 1e-3, 0.5, 0.75, 1.0 - 0x1p-53, 1.0, 0.9, 1e-20};
    const double epsilons[] = {1e-40, DBL_EPSILON, 1e-9, 0.0, 1e-300};
    long bad = 0, ndiv = 0;
    for (long j = 0; j < n; ++j) {
        const double thres = thresholds[j % 8];
        const double eps = epsilons[(j / 8) % 5];
        const double gap = sg::conf_gap(thres, eps);
        double ppos, pneg;
        const int kind = (int)(next() % 8);
        const double s = std::ldexp(0.5 + 0.5 * u01(), -(int)(next() % 140));
        if (kind <= 3) {
            // quotient near thres (kind 0-1) or near the midpoint above it (2-3):
            // ppos = q s for q a few ulps from the target, pneg = s - ppos
            const double ulp = std::nextafter(thres, 2.0) - thres;
            const double target = kind <= 1 ? thres : thres + 0.5 * ulp;
            const double q = target + ((double)(int64_t)(next() % 9) - 4.0) * 0.25 * ulp;
            ppos = q * s;
            pneg = s - ppos;
            if (kind == 1) pneg = std::nextafter(pneg, (next() & 1) ? 1.0 : -1.0);
        } else if (kind == 4) {
            // a tie: P_pos / s exactly the midpoint (s a power of two times an
            // odd number of the midpoint's precision)
            const double ulp = std::nextafter(thres, 2.0) - thres;
            const double sp = std::ldexp(1.0, -(int)(next() % 100));
            ppos = (thres + 0.5 * ulp) * sp; // rounds to one of its neighbours
            pneg = sp - ppos;
        } else if (kind == 5) {
            // P_notGuess cancelled to zero or below: P_neg <= 0
            ppos = u01() * s;
            pneg = -u01() * ldexp(s, -(int)(next() % 60));
        } else if (kind == 6) {
            // around the epsilon cut
            const double e = eps > 0 ? eps : 1e-300;
            ppos = e * u01();
            pneg = e * (u01() - 0.25);
        } else {
            // special values
            const double sp[] = {0.0, -0.0, INFINITY, -INFINITY, NAN, 1.0, DBL_MIN, DBL_MAX};
            ppos = (next() & 1) ? sp[next() % 8] : u01();
            pneg = (next() & 1) ? sp[next() % 8] : u01() * 0x1p-8;
        }
        bad += sg::conf_above(ppos, pneg, eps, thres, gap) != reference(ppos, pneg, eps, thres);
        ndiv += divides(ppos, pneg, eps, thres, gap);
    }
    printf("conf %ld %ld %ld\n", n, bad, ndiv);
    return 0;
}
