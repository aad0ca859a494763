// Host builds of the decoder's logarithm and exponential (sogrand_amd/csrc/
// sg_log.h and sg_exp.h, the same sources the kernels compile) compared bit for
// bit with the C library's log and exp on arguments of the shapes the decoder
// forms, and on special values. Prints two lines:
//   log n mismatches special_fail
//   exp n mismatches special_fail
// Usage: log_check [n] [seed]
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "sg_exp.h"
#include "sg_log.h"

static uint64_t bits(double x)
{
    uint64_t u;
    memcpy(&u, &x, 8);
    return u;
}

static bool same(double a, double b)
{
    return bits(a) == bits(b) || (std::isnan(a) && std::isnan(b));
}

int main(int argc, char **argv)
{
    const long n = argc > 1 ? atol(argv[1]) : 200000;
    uint64_t s = argc > 2 ? strtoull(argv[2], nullptr, 10) : 88172645463325252ull;
    auto u01 = [&]() {
        s ^= s << 13;
        s ^= s >> 7;
        s ^= s << 17;
        return (double)(s >> 11) * 0x1p-53;
    };
    sg::ExpEntry tab[sg::kExpTableEntries];
    for (int i = 0; i < sg::kExpTableEntries; ++i) tab[i] = sg::exp_table_entry(i);

    long lbad = 0;
    for (long j = 0; j < n; ++j) {
        double x;
        switch (j % 6) {
        case 0: x = 1.0 + (u01() - 0.5) * 0.2; break;            // both sides of 1
        case 1: x = 1.0 + u01(); break;                           // (1, 2]
        case 2: x = exp((u01() - 0.5) * 200.0); break;            // e^-100 .. e^100
        case 3: x = 1.0 + exp(-50.0 * u01()); break;              // JacLog(-a): 1 + exp(-a)
        case 4: x = 1.0 + exp(50.0 * u01()); break;               // JacLog(a): 1 + exp(a)
        default: x = 1e-40 + u01() * ldexp(1.0, -(int)(j % 140)); // probability sums
        }
        lbad += !same(sg::log_pos(x, sg::kLogTable), log(x));
    }
    long lspecial = 0;
    const double lsp[] = {1.0, 0.0, -0.0, -1.0, INFINITY, -INFINITY, NAN, 0x1p-1074, 0x1.8p-1070,
                          0x1p-1022, DBL_MAX, 1.0 - 0x1p-4, 1.0 + 0x1.09p-4, 0x1.0ffffffffffffp0,
                          0x1.6p-1, 0x1.6p0, 1.0 - 0x1p-53, 1.0 + 0x1p-52};
    for (double x : lsp) lspecial += !same(sg::log_pos(x, sg::kLogTable), log(x));
    for (int k = -1074; k <= 1023; ++k) lspecial += !same(sg::log_pos(ldexp(1.0, k), sg::kLogTable),
                                                          log(ldexp(1.0, k)));
    printf("log %ld %ld %ld\n", n, lbad, lspecial);

    long ebad = 0;
    for (long j = 0; j < n; ++j) {
        double x;
        switch (j % 5) {
        case 0: x = (u01() - 0.5) * 100.0; break;   // JacLog's exp(+-a), |a| <= 50
        case 1: x = -60.0 * u01(); break;           // guess weights exp(-pm), the usual range
        case 2: x = -800.0 * u01(); break;          // guess weights, large metrics
        case 3: x = (u01() - 0.5) * 1500.0; break;  // the soft output's exp(x), any x
        default: x = -745.2 + 40.0 * u01(); break;  // subnormal results
        }
        ebad += !same(sg::exp_of(x, tab), exp(x));
    }
    long especial = 0;
    const double esp[] = {0.0, -0.0, 0x1p-60, -0x1p-60, 0x1p-1074, 512.0, -512.0, 709.78, 709.79,
                          -708.39, -745.13, -745.14, 1024.0, -1024.0, INFINITY, -INFINITY, NAN,
                          DBL_MAX, -DBL_MAX};
    for (double x : esp) especial += !same(sg::exp_of(x, tab), exp(x));
    for (int k = -1000; k <= 1000; ++k) especial += !same(sg::exp_of(k * 0.6931471805599453, tab),
                                                          exp(k * 0.6931471805599453));
    printf("exp %ld %ld %ld\n", n, ebad, especial);
    return 0;
}
