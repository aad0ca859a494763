// Host build of the decoder's logarithm (sogrand_amd/csrc/sg_log.h, the same
// source the kernels compile) measured against the C library's log and
// against logl rounded to double. Prints one line of counts:
//   n mismatch_vs_libc mismatch_vs_logl libc_vs_logl max_err_ulp_x1e4 special_fail
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>

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
    printf("%ld %ld %ld %ld %ld %ld\n", n, vs_libc, vs_logl, libc_vs_logl, (long)(max_err * 1e4),
           special);
    return 0;
}
