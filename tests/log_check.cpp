// Host builds of the decoder's logarithm and exponential (sogrand_amd/csrc/
// sg_log.h and sg_exp.h, the same sources the kernels compile) measured against
// the C library's log and exp and against logl / expl rounded to double.
// Prints two lines of counts, "log ..." and "exp ...":
//   n mismatch_vs_libc mismatch_vs_long_double libc_vs_long_double max_err_ulp_x1e4 special_fail
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>

#include "sg_exp.h"
#include "sg_log.h"

int main(int argc, char **argv)
{
    const long n = argc > 1 ? atol(argv[1]) : 200000;
    std::mt19937_64 g(12345);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    long vs_libc = 0, vs_logl = 0, libc_vs_logl = 0;
    double max_err = 0.0;
    auto one = [&](double x) {
        const double a = sg::log_pos(x, sg::kLogTable);
        const long double L = logl((long double)x);
        const double b = log(x), cr = (double)L;
        vs_libc += a != b;
        vs_logl += a != cr;
        libc_vs_logl += b != cr;
        const long double ulp = (long double)nextafter(fabs(cr), INFINITY) - fabsl((long double)cr);
        const double e = (double)(fabsl((long double)a - L) / ulp);
        if (e > max_err) max_err = e;
    };
    for (long j = 0; j < n; ++j) {
        double x;
        switch (j % 5) {
        case 0: x = 1.0 + (u(g) - 0.5) * 0.02; break;   // near 1, both sides
        case 1: x = 1.0 + u(g); break;                  // (1, 2]
        case 2: x = exp((u(g) - 0.5) * 200.0); break;   // e^-100 .. e^100
        case 3: x = 1.0 + exp(-40.0 * u(g)); break;     // 1 + exp(-a), the set-up's argument
        default: x = 1e-40 + u(g); break;               // probability sums
        }
        one(x);
    }
    // values the arithmetic fixes exactly
    long special = 0;
    special += sg::log_pos(1.0, sg::kLogTable) != 0.0;
    for (int k = -1000; k <= 1000; k += 37)
        special += sg::log_pos(ldexp(1.0, k), sg::kLogTable) != (double)(k * logl(2.0L));
    printf("log %ld %ld %ld %ld %ld %ld\n", n, vs_libc, vs_logl, libc_vs_logl,
           (long)(max_err * 1e4), special);

    sg::ExpEntry tab[sg::kExpTableEntries];
    for (int i = 0; i < sg::kExpTableEntries; ++i) tab[i] = sg::exp_table_entry(i);
    long e_libc = 0, e_ld = 0, libc_ld = 0;
    double e_max = 0.0;
    for (long j = 0; j < n; ++j) {
        double x;
        switch (j % 3) {
        case 0: x = (u(g) - 0.5) * 100.0; break;  // JacLog's exp(+-a), |a| <= 50
        case 1: x = -700.0 * u(g); break;         // guess weights exp(-pm)
        default: x = (u(g) - 0.5) * 1400.0; break; // soft output's exp(x)
        }
        const double a = sg::exp_of(x, tab);
        const long double E = expl((long double)x);
        const double b = exp(x), cr = (double)E;
        e_libc += a != b;
        e_ld += a != cr;
        libc_ld += b != cr;
        const long double ulp = (long double)nextafter(cr, INFINITY) - (long double)cr;
        const double e = (double)(fabsl((long double)a - E) / ulp);
        if (e > e_max) e_max = e;
    }
    long e_special = 0;
    e_special += sg::exp_of(0.0, tab) != 1.0;
    e_special += sg::exp_of(-746.0, tab) != 0.0;
    e_special += sg::exp_of(710.0, tab) != INFINITY;
    for (int k = -1000; k <= 1000; k += 37) // x = k ln2 rounded: within 2^-52 |k| of 2^k
        e_special += fabs(sg::exp_of((double)(k * logl(2.0L)), tab) / ldexp(1.0, k) - 1.0) > 0x1p-40;
    printf("exp %ld %ld %ld %ld %ld %ld\n", n, e_libc, e_ld, libc_ld, (long)(e_max * 1e4),
           e_special);
    return 0;
}
