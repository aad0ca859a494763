/* Drives the C oracle (oracle/sogrand_oracle.c) through its whole API for the
 * AddressSanitizer / UndefinedBehaviorSanitizer build of tests/test_sanitizers.py:
 * code tables, the tie-ordered sort (ties, infinities, NaN), schedule
 * prefixes, component decodes of ordinary, tied, non-finite and huge lines,
 * and whole-cube decodes with soft outputs, conditioning and statistics on two
 * threads. Prints "ok" and exits 0 when every call returned as expected. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sogrand_oracle.h"

static uint64_t rng_state = 88172645463325252ull;

static double u01(void)
{
    rng_state ^= rng_state << 13;
    rng_state ^= rng_state >> 7;
    rng_state ^= rng_state << 17;
    return (double)(rng_state >> 11) * 0x1p-53;
}

static double gauss(void)
{
    double u1 = u01();
    if (u1 <= 0.0) u1 = 0x1p-53;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u01());
}

int main(void)
{
    uint8_t G[128], H[128], hcol[16];
    og_code_tables(G, H, hcol);

    for (int k = 0; k < 200; k++) {
        double a[16];
        uint8_t perm[16];
        for (int i = 0; i < 16; i++) {
            a[i] = (double)(int)(4.0 * u01());
            perm[i] = (uint8_t)i;
        }
        if (k % 5 == 1) a[(k * 7) % 16] = INFINITY;
        if (k % 5 == 2) a[(k * 3) % 16] = NAN;
        og_sort_with_perm(a, perm, 16);
    }

    uint16_t masks[4096];
    const int32_t ics[] = {0, 1, 7, 55, 56, 1000};
    for (size_t j = 0; j < sizeof ics / sizeof ics[0]; j++)
        if (og_schedule_prefix(ics[j], masks, 4096) != 4096) return 1;

    double la[16], le[16], x[16];
    og_line_trace tr;
    for (int k = 0; k < 120; k++) {
        for (int i = 0; i < 16; i++) x[i] = 0.8 + 1.3 * gauss();
        if (k % 6 == 1)
            for (int i = 0; i < 16; i++) x[i] = (double)(int)(3.0 * x[i]);
        if (k % 6 == 2) x[k % 16] = (k & 8) ? INFINITY : -INFINITY;
        if (k % 6 == 3) x[k % 16] = NAN;
        if (k % 6 == 4)
            for (int i = 0; i < 16; i++) x[i] *= 300.0;
        if (og_siso(x, 3, k % 4 == 0 ? 100 : UINT64_MAX, 1.0 - 1e-5, la, le, &tr) != 0) return 2;
    }

    /* three cubes of the all-zero codeword at 3 dB (sigma^2 = 2.0047), plus
     * one with an infinite entry */
    enum { NC = 4 };
    double *llr = malloc(sizeof(double) * NC * OG_CUBE);
    double *soft = malloc(sizeof(double) * NC * OG_CUBE);
    double *cond = malloc(sizeof(double) * NC * OG_CUBE);
    uint8_t bits[NC * 64];
    og_cw_stats st[NC];
    if (!llr || !soft || !cond) return 3;
    const double s2 = 2.0047;
    for (int i = 0; i < NC * OG_CUBE; i++) llr[i] = 2.0 * (1.0 + sqrt(s2) * gauss()) / s2;
    llr[3 * OG_CUBE + 77] = INFINITY;
    if (og_decode_cubes_cond(llr, NC, 3, 30, UINT64_MAX, 1e-5, 0.7, bits, soft, cond, st, 2) != 0)
        return 4;
    if (og_decode_cubes(llr, 2, 3, 2, 1000, 1e-5, 0.7, bits, NULL, NULL, 1) != 0) return 5;
    (void)og_cube_is_codeword(soft);
    free(llr);
    free(soft);
    free(cond);
    printf("ok\n");
    return 0;
}
