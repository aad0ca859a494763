/*
 * CPU oracle for the SOGRAND (16,8)^3 cubic product-code decoder.
 *
 * TEST INFRASTRUCTURE ONLY: the checker the HIP decoder is compared against,
 * never part of the product. It restates the semantics of the reference CPU
 * decoder SOGRAND_C/cubic_decoder_final.c (all citations below are to that
 * file unless stated) in this repository's own formulation:
 *
 *   - error patterns are 16-bit masks over sorted positions, taken from one
 *     table of all 65535 nonzero masks ordered by (Hamming weight w, position
 *     sum S, lexicographic order of the position tuple);
 *   - the schedule for intercept IC visits the (w, S) levels in order of
 *     logistic weight wt = IC*w + S, then w, as the landslide loop does
 *     (:642-744); the lexicographic order inside a level is the order the
 *     reference's mountain_build/landslide step produces (:689-734, :775-789);
 *   - parity checks are 8-bit syndromes: s0 of the hard decision XOR the
 *     H-columns of the flipped positions (equivalent to ParityCheck, :748-755);
 *   - path metrics use per-position increments delta_b = JacLog(a_b) -
 *     JacLog(-a_b), added in ascending sorted position, exactly the sums getPM
 *     forms (:765), so every double is rounded as in the reference;
 *   - early termination is "every one of the 768 lines of the hard decision is
 *     a component codeword", equivalent to the re-encode test (:317-391).
 *
 * Transcendentals are glibc's exp/log, as in the reference. Compiled with
 * -ffp-contract=off: no multiply-add is fused.
 *
 * Two process-wide switches restate the reference's other decoders:
 * og_set_epsilon(DBL_EPSILON) gives SOGRAND_C/cubic_decoder_updated.c, and
 * og_set_variant(1) the earlier SOGRAND_C/cubic_decoder.c (JacLog cut at +-30,
 * IC guard, P_notGuess never lowered, P_neg of the last guess under a budget,
 * clamps 1e-9 and 1e-30; see og_siso).
 */
#define _GNU_SOURCE
#include "sogrand_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

/* ------------------------------------------------------------------------- */
/* Code tables: (16,8) systematic CRC code, Koopman polynomial 0xeb           */
/* (getGH_sys_CRC / koopman2matlab, :454-519).                                */
/* ------------------------------------------------------------------------- */

/* Koopman 0xeb drops the x^0 term: g(x) = (0xeb << 1) | 1 = x^8+x^7+x^6+x^4+x^2+x+1. */
#define OG_GPOLY 0x1d7u

static uint8_t og_G[8][16], og_H[8][16], og_hcol[16];

/* Remainder of x^(15-i) mod g(x): bit (7-j) of it is P[i][j] (:491-503). */
static unsigned og_parity_row(int i)
{
    unsigned v = 1u << (15 - i);
    for (int deg = 15; deg >= 8; deg--)
        if (v & (1u << deg)) v ^= OG_GPOLY << (deg - 8);
    return v;
}

static void og_build_code(void)
{
    memset(og_G, 0, sizeof og_G);
    memset(og_H, 0, sizeof og_H);
    for (int i = 0; i < 8; i++) {
        unsigned rem = og_parity_row(i);
        og_G[i][i] = 1; /* G = [I | P] (:505-508) */
        for (int j = 0; j < 8; j++) {
            uint8_t pij = (uint8_t)((rem >> (7 - j)) & 1u);
            og_G[i][8 + j] = pij;
            og_H[j][i] = pij; /* H = [P^T | I] (:510-513) */
        }
        og_H[i][8 + i] = 1;
    }
    for (int c = 0; c < 16; c++) {
        uint8_t s = 0;
        for (int r = 0; r < 8; r++) s |= (uint8_t)(og_H[r][c] << r);
        og_hcol[c] = s;
    }
}

/* ------------------------------------------------------------------------- */
/* Pattern levels: all nonzero 16-bit masks grouped by (w, S), S = sum of     */
/* (position + 1); lexicographic order of the ascending position tuple        */
/* inside each group. Lexicographic ascending == bit-reversed mask descending */
/* (the lowest differing position decides, and its owner comes first).        */
/* ------------------------------------------------------------------------- */

#define OG_SMAX 136
static uint16_t og_masks[65535];
static int32_t og_lvl_start[17][OG_SMAX + 2]; /* [w][S] -> first index; [w][S+1] bounds it */
static int32_t og_lvl_count[17][OG_SMAX + 1];

static pthread_once_t og_once = PTHREAD_ONCE_INIT;

static unsigned og_rev16(unsigned m)
{
    unsigned r = 0;
    for (int b = 0; b < 16; b++)
        if (m & (1u << b)) r |= 1u << (15 - b);
    return r;
}

static int og_possum(unsigned m)
{
    int s = 0;
    for (int b = 0; b < 16; b++)
        if (m & (1u << b)) s += b + 1;
    return s;
}

static void og_build_levels(void)
{
    memset(og_lvl_count, 0, sizeof og_lvl_count);
    for (unsigned m = 1; m < 65536u; m++) og_lvl_count[__builtin_popcount(m)][og_possum(m)]++;
    int32_t off = 0;
    for (int w = 0; w <= 16; w++)
        for (int s = 0; s <= OG_SMAX + 1; s++) {
            og_lvl_start[w][s] = off;
            if (s <= OG_SMAX) off += og_lvl_count[w][s];
        }
    int32_t fill[17][OG_SMAX + 1];
    memset(fill, 0, sizeof fill);
    /* visiting bit-reversed values in descending order appends each group in
     * lexicographic order */
    for (int r = 65535; r >= 1; r--) {
        unsigned m = og_rev16((unsigned)r);
        int w = __builtin_popcount(m), s = og_possum(m);
        og_masks[og_lvl_start[w][s] + fill[w][s]++] = (uint16_t)m;
    }
}

static void og_init_once(void)
{
    og_build_code();
    og_build_levels();
}

static void og_init(void) { pthread_once(&og_once, og_init_once); }

void og_code_tables(uint8_t *G, uint8_t *H, uint8_t *hcol)
{
    og_init();
    if (G) memcpy(G, og_G, sizeof og_G);
    if (H) memcpy(H, og_H, sizeof og_H);
    if (hcol) memcpy(hcol, og_hcol, sizeof og_hcol);
}

/* Smallest / largest position sum of a w-subset of {1..16}. */
static int64_t og_smin(int w) { return (int64_t)w * (w + 1) / 2; }
static int64_t og_smax(int w) { return (int64_t)w * (33 - w) / 2; }

/* Walk of the (wt, w) levels in schedule order (:645-744). wt, ic in 64-bit. */
typedef struct {
    int64_t ic, wt;
    int w;
} og_walk;

/* First level with (wt, w) strictly after the current one; 0 when exhausted. */
static int og_walk_next(og_walk *s)
{
    for (int w = s->w + 1; w <= 16; w++) {
        int64_t S = s->wt - s->ic * w;
        if (S >= og_smin(w) && S <= og_smax(w)) {
            s->w = w;
            return 1;
        }
    }
    int64_t best = -1;
    for (int w = 1; w <= 16; w++) {
        int64_t lo = s->ic * w + og_smin(w), hi = s->ic * w + og_smax(w);
        int64_t cand = s->wt + 1 > lo ? s->wt + 1 : lo;
        if (cand <= hi && (best < 0 || cand < best)) best = cand;
    }
    if (best < 0) return 0;
    s->wt = best;
    for (int w = 1; w <= 16; w++) {
        int64_t S = best - s->ic * w;
        if (S >= og_smin(w) && S <= og_smax(w)) {
            s->w = w;
            return 1;
        }
    }
    return 0;
}

static void og_walk_begin(og_walk *s, int32_t ic)
{
    s->ic = ic;
    s->wt = ic; /* the first level searched is wt = IC + 1 (:645) */
    s->w = 16;
}

int og_schedule_prefix(int32_t ic, uint16_t *masks, int count)
{
    og_init();
    og_walk s;
    og_walk_begin(&s, ic);
    int n = 0;
    while (n < count && og_walk_next(&s)) {
        int S = (int)(s.wt - s.ic * s.w);
        int32_t a = og_lvl_start[s.w][S], e = a + og_lvl_count[s.w][S];
        for (int32_t i = a; i < e && n < count; i++) masks[n++] = og_masks[i];
    }
    return n;
}

/* ------------------------------------------------------------------------- */
/* Line decode (SOGRAND_bitSO :521-585 and sogrand_main_logic :587-746).      */
/* ------------------------------------------------------------------------- */

/* :762; the earlier cubic_decoder.c:699 cuts at +-30 */
static double og_jaclog(double x, double lim)
{
    if (x > lim) return x;
    if (x < -lim) return 0.0;
    return log(1.0 + exp(x));
}

/* Ascending sort carrying a permutation. The reference sorts with a Hoare
 * partition around the middle element (:763); ties among equal magnitudes are
 * left in the order that procedure produces, so the same partition steps are
 * used here (iteratively, with an explicit range stack). */
void og_sort_with_perm(double *a, uint8_t *perm, int n)
{
    int lo_stack[64], len_stack[64], top = 0;
    lo_stack[top] = 0;
    len_stack[top++] = n;
    while (top > 0) {
        top--;
        int lo = lo_stack[top], len = len_stack[top];
        if (len < 2) continue;
        double *x = a + lo;
        uint8_t *p = perm + lo;
        double pivot = x[len / 2];
        int i = 0, j = len - 1;
        while (i <= j) {
            while (x[i] < pivot) i++;
            while (x[j] > pivot) j--;
            if (i <= j) {
                double t = x[i];
                x[i] = x[j];
                x[j] = t;
                uint8_t tp = p[i];
                p[i] = p[j];
                p[j] = tp;
                i++;
                j--;
            }
        }
        /* the reference recurses on [0, j] first, then [i, len): push in reverse */
        if (i < len) {
            lo_stack[top] = lo + i;
            len_stack[top++] = len - i;
        }
        if (j > 0) {
            lo_stack[top] = lo;
            len_stack[top++] = j + 1;
        }
    }
}

/* x86-64 cvttsd2si to a 32-bit register: NaN or out of range gives INT32_MIN. */
static int32_t og_cvtt_i32(double r)
{
    if (r >= -2147483648.0 && r < 2147483648.0) return (int32_t)r;
    return INT32_MIN;
}

/* EPSILON of the soft output and the confidence test: 1e-40 in
 * cubic_decoder_final.c:16, DBL_EPSILON in the reference's cubic_decoder_updated.c:16. */
static double og_epsilon = 1e-40;

void og_set_epsilon(double eps) { og_epsilon = eps; }

static int og_variant = 0;

void og_set_variant(int variant) { og_variant = variant; }

int og_siso(const double *llr, int L, uint64_t Tmax, double thres, double *L_APP, double *L_E,
            og_line_trace *trace)
{
    og_init();
    const int legacy = og_variant == 1;
    const double eps = og_epsilon; /* EPSILON (:16) */
    /* cubic_decoder.c: 1e-9 in getLConf (:710), the P_notinlist floor and the
     * clamps (:503, :509-510); 1e-30 inside the logs (:533) */
    const double eps_conf = legacy ? 1e-9 : eps;
    const double eps_clamp = legacy ? 1e-9 : eps;
    const double eps_log = legacy ? 1e-30 : eps;
    const double jlim = legacy ? 30.0 : 50.0;
    if (L < 1) L = 1;
    if (L > OG_MAXL) L = OG_MAXL;
    if (Tmax == 0) Tmax = 0x7fffffff; /* :607 with Inf from :14 */

    double a[16];
    uint8_t perm[16];
    unsigned hd = 0;
    uint8_t s0 = 0;
    for (int i = 0; i < 16; i++) {
        if (!(llr[i] > 0.0)) { /* HardDec (:758) */
            hd |= 1u << i;
            s0 ^= og_hcol[i];
        }
        a[i] = fabs(llr[i]);
        perm[i] = (uint8_t)i;
    }
    /* PM of the hard decision, natural order, before sorting (:617) */
    double pm_hd = 0.0;
    for (int i = 0; i < 16; i++) pm_hd += og_jaclog(-a[i], jlim);
    og_sort_with_perm(a, perm, 16);

    /* IC (:619-622) */
    double beta = (a[7] - a[0]) / 7.0;
    int32_t ic = og_cvtt_i32(round(a[0] / beta - 1.0));
    if (ic < 0) ic = 0;
    if (legacy && !(beta > 1e-9)) ic = 0; /* cubic_decoder.c:580-585 */

    double delta[16];
    uint8_t hp[16];
    for (int b = 0; b < 16; b++) {
        delta[b] = og_jaclog(a[b], jlim) - og_jaclog(-a[b], jlim);
        hp[b] = og_hcol[perm[b]];
    }

    int nl = 0;
    double lpm[OG_MAXL];
    uint16_t lcw[OG_MAXL];
    double pnl = 0.0;
    double pnot = 1.0; /* P_notGuess, even = 0 (:614) */
    uint64_t T = 0;
    if (trace) {
        memset(trace, 0, sizeof *trace);
        trace->ic = ic;
        memcpy(trace->perm, perm, 16);
    }

    /* Record a codeword hit; returns 1 if the search stops (:628-641, :674-686). */
#define OG_HIT(PM, CWMASK, WT)                                                           \
    do {                                                                                 \
        lpm[nl] = (PM);                                                                  \
        lcw[nl] = (uint16_t)(CWMASK);                                                    \
        double ppos = 0.0;                                                               \
        for (int q = 0; q <= nl; q++) ppos += exp(-lpm[q]);                              \
        double pneg = 0.00390625 * pnot; /* 2^-(n-k) (:771) */                           \
        pnl = pneg;                                                                      \
        double conf = (ppos + pneg > eps_conf) ? ppos / (ppos + pneg) : 1.0;             \
        if (trace) {                                                                     \
            trace->pm[nl] = lpm[nl];                                                     \
            trace->conf[nl] = conf;                                                      \
            trace->t_hit[nl] = (uint32_t)T;                                              \
            trace->wt_hit[nl] = (int32_t)(WT);                                           \
            trace->cw[nl] = lcw[nl];                                                     \
        }                                                                                \
        nl++;                                                                            \
        if (conf > thres || nl == L) goto search_done;                                   \
    } while (0)

    /* guess 1: the hard decision itself (:624-641) */
    T = 1;
    if (legacy)
        pnl = exp(-pm_hd); /* cubic_decoder.c:590, overwritten by getLConf at a hit */
    else
        pnot -= exp(-pm_hd);
    if (s0 == 0) OG_HIT(pm_hd, hd, 0);

    {
        og_walk s;
        og_walk_begin(&s, ic);
        int64_t cur_wt = -1;
        while (og_walk_next(&s)) {
            if (s.wt != cur_wt) { /* loop guard once per wt level (:652) */
                if (!(T < Tmax)) break;
                cur_wt = s.wt;
            }
            int S = (int)(s.wt - s.ic * s.w);
            int32_t i0 = og_lvl_start[s.w][S], i1 = i0 + og_lvl_count[s.w][S];
            for (int32_t idx = i0; idx < i1; idx++) {
                unsigned m = og_masks[idx];
                double pm = pm_hd;
                uint8_t syn = s0;
                unsigned cw = hd;
                for (unsigned r = m; r; r &= r - 1) {
                    int b = __builtin_ctz(r);
                    pm += delta[b];
                    syn ^= hp[b];
                    cw ^= 1u << perm[b];
                }
                T++;
                if (legacy)
                    pnl = exp(-pm); /* cubic_decoder.c:638 */
                else
                    pnot -= exp(-pm);
                if (syn == 0) OG_HIT(pm, cw, s.wt);
            }
        }
    }
#undef OG_HIT

search_done:
    if (trace) {
        trace->n_list = nl;
        trace->T = (uint32_t)T;
        trace->pnl = pnl;
    }
    if (!L_APP && !L_E) return 0;

    /* Soft output (:536-575). */
    double la[16];
    if (nl == 0) {
        for (int i = 0; i < 16; i++) la[i] = llr[i];
        if (L_APP) memcpy(L_APP, la, sizeof la);
        if (L_E)
            for (int i = 0; i < 16; i++) L_E[i] = 0.0;
        return 0;
    }
    double pnotin = fmax(pnl, eps_clamp);
    double p[OG_MAXL];
    for (int j = 0; j < nl; j++) p[j] = exp(-lpm[j]);
    for (int i = 0; i < 16; i++) {
        double pp1 = 1.0 / (1.0 + exp(llr[i]));
        double pp0 = 1.0 - pp1;
        pp1 = fmin(fmax(pp1, eps_clamp), 1.0 - eps_clamp);
        pp0 = fmin(fmax(pp0, eps_clamp), 1.0 - eps_clamp);
        double p1 = 0.0, p0 = 0.0;
        for (int j = 0; j < nl; j++) {
            double c = (double)((lcw[j] >> i) & 1u);
            p1 += c * p[j];
            p0 += (1.0 - c) * p[j];
        }
        p0 += pnotin * pp0;
        p1 += pnotin * pp1;
        la[i] = log(fmax(p0, eps_log)) - log(fmax(p1, eps_log));
    }
    if (L_APP) memcpy(L_APP, la, sizeof la);
    if (L_E)
        for (int i = 0; i < 16; i++) L_E[i] = la[i] - llr[i];
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Cube level.                                                                */
/* ------------------------------------------------------------------------- */

static int og_line_syndrome(unsigned word)
{
    uint8_t s = 0;
    for (unsigned r = word; r; r &= r - 1) s ^= og_hcol[__builtin_ctz(r)];
    return s;
}

/* Line geometry of the product code of dimension dim (2 or 3): positions
 * a0 + 16 a1 (+ 256 a2); line (x, y) of axis ax covers base + j stride[ax],
 * base = x o1[ax] + y o2[ax], with 16^(dim-1) lines per axis (y = 0 for dim 2).
 * The square code's axes are the cube's first two. */
static const int og_stride[3] = {1, 16, 256};
static const int og_o1[2][3] = {{16, 1, 0}, {16, 1, 1}};
static const int og_o2[3] = {256, 256, 16};

static int og_is_product_codeword(const double *L_APP, int dim)
{
    og_init();
    const int *o1 = og_o1[dim - 2], ny = dim == 3 ? 16 : 1;
    for (int ax = 0; ax < dim; ax++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < 16; x++) {
                int base = x * o1[ax] + y * og_o2[ax];
                unsigned w = 0;
                for (int j = 0; j < 16; j++)
                    if (!(L_APP[base + j * og_stride[ax]] > 0.0)) w |= 1u << j;
                if (og_line_syndrome(w)) return 0;
            }
    return 1;
}

int og_cube_is_codeword(const double *L_APP) { return og_is_product_codeword(L_APP, 3); }

int og_square_is_codeword(const double *L_APP) { return og_is_product_codeword(L_APP, 2); }

typedef struct {
    const double *llr;
    int dim; /* 2: (16,8)^2 square code, 3: (16,8)^3 cubic code */
    size_t n_cw;
    int L, Imax;
    uint64_t Tmax;
    double thres, alpha;
    uint8_t *bits;
    double *soft;
    double *cond;
    og_cw_stats *stats;
    size_t next; /* shared work counter */
    int bad;
    pthread_mutex_t mu;
} og_job;

/* Rounding-error bound, relative, of the reference's P_notGuess at the last list
 * insertion of a line: T guesses of magnitude <= 1 subtracted from 1.0, each
 * with increments whose own error is a few ulps (T + 64) * 2^-50 in absolute
 * terms, divided by the value. 0 for an empty list (L_APP = input exactly). */
static double og_chain_cond(const og_line_trace *tr)
{
    if (tr->n_list == 0) return 0.0;
    double P = 256.0 * tr->pnl;
    if (!(fabs(P) > 0.0)) return INFINITY;
    return ((double)tr->T + 64.0) * 8.8817841970012523e-16 / fabs(P); /* 2^-50 */
}

/* One codeword of the product code of dimension dim: the reference's
 * per-codeword loop (final.c:119-289) with one phase per axis; the square code
 * runs the same loop over its two axes (the reference ships no square decoder:
 * its channel simulator accepts dimension 2, channel_sim.c:26-29). */
static int og_decode_one(int dim, const double *Lch, int L, int Imax, uint64_t Tmax,
                         double thres, double alpha, uint8_t *bits, double *soft, double *cond,
                         og_cw_stats *st)
{
    const int npos = dim == 3 ? OG_CUBE : 256, ny = dim == 3 ? 16 : 1;
    const int *o1 = og_o1[dim - 2];
    const int *stride = og_stride, *o2 = og_o2;
    double *LE = (double *)calloc((size_t)npos, sizeof(double));
    double *LA = (double *)malloc((size_t)npos * sizeof(double));
    double *in = (double *)malloc((size_t)npos * sizeof(double));
    int rc = 0;
    uint32_t halves = 0, conv = 0;
    uint64_t ng = 0, ngp = 0;
    memcpy(LA, Lch, (size_t)npos * sizeof(double));
    for (int it = 1; it <= Imax && !conv; it++) {
        for (int ax = 0; ax < dim; ax++) {
            halves++; /* n_iter += 0.5 before the phase (:145,:184,:223) */
            for (int i = 0; i < npos; i++) {
                double t = alpha * LE[i]; /* L_A, then input, two roundings (:148-155) */
                in[i] = Lch[i] + t;
            }
            uint32_t gmax = 0;
            for (int y = 0; y < ny; y++)
                for (int x = 0; x < 16; x++) {
                    int base = x * o1[ax] + y * o2[ax];
                    double v[16], la[16], le[16];
                    og_line_trace tr;
                    for (int j = 0; j < 16; j++) v[j] = in[base + j * stride[ax]];
                    if (og_siso(v, L, Tmax, thres, la, le, &tr)) rc = -1;
                    ng += tr.T;
                    if (tr.T > gmax) gmax = tr.T;
                    const double kc = og_chain_cond(&tr);
                    for (int j = 0; j < 16; j++) {
                        LA[base + j * stride[ax]] = la[j];
                        LE[base + j * stride[ax]] = le[j];
                        if (cond) cond[base + j * stride[ax]] = kc;
                    }
                }
            ngp += gmax;
            if (og_is_product_codeword(LA, dim)) {
                conv = 1;
                break;
            }
        }
    }
    if (bits) {
        /* info bit b = a0 + 8 a1 (+ 64 a2), MSB first (:267-280) */
        const int ninfo = dim == 3 ? OG_INFO : 64;
        memset(bits, 0, (size_t)ninfo / 8);
        for (int b = 0; b < ninfo; b++) {
            int a0 = b & 7, a1 = (b >> 3) & 7, a2 = b >> 6;
            if (!(LA[a0 + 16 * a1 + 256 * a2] > 0.0)) bits[b >> 3] |= (uint8_t)(0x80u >> (b & 7));
        }
    }
    if (soft) memcpy(soft, LA, (size_t)npos * sizeof(double));
    if (st) {
        st->half_iters = halves;
        st->converged = conv;
        st->NG = ng;
        st->NG_p = ngp;
    }
    free(LE);
    free(LA);
    free(in);
    return rc;
}

static void *og_worker(void *arg)
{
    og_job *J = (og_job *)arg;
    for (;;) {
        pthread_mutex_lock(&J->mu);
        size_t c = J->next++;
        pthread_mutex_unlock(&J->mu);
        if (c >= J->n_cw) break;
        const size_t npos = J->dim == 3 ? OG_CUBE : 256, nbytes = J->dim == 3 ? 64 : 8;
        int rc = og_decode_one(J->dim, J->llr + c * npos, J->L, J->Imax, J->Tmax, J->thres,
                               J->alpha, J->bits ? J->bits + c * nbytes : NULL,
                               J->soft ? J->soft + c * npos : NULL,
                               J->cond ? J->cond + c * npos : NULL,
                               J->stats ? J->stats + c : NULL);
        if (rc) {
            pthread_mutex_lock(&J->mu);
            J->bad = 1;
            pthread_mutex_unlock(&J->mu);
        }
    }
    return NULL;
}

int og_decode_cubes(const double *llr, size_t n_cw, int L, int Imax, uint64_t Tmax, double p_et,
                    double alpha, uint8_t *bits, double *soft, og_cw_stats *stats, int threads)
{
    return og_decode_cubes_cond(llr, n_cw, L, Imax, Tmax, p_et, alpha, bits, soft, NULL, stats,
                                threads);
}

static int og_decode_product(int dim, const double *llr, size_t n_cw, int L, int Imax,
                             uint64_t Tmax, double p_et, double alpha, uint8_t *bits,
                             double *soft, double *cond, og_cw_stats *stats, int threads)
{
    og_init();
    og_job J;
    J.llr = llr;
    J.dim = dim;
    J.n_cw = n_cw;
    J.L = L;
    J.Imax = Imax;
    J.Tmax = Tmax;
    J.thres = 1.0 - p_et;
    J.alpha = alpha;
    J.bits = bits;
    J.soft = soft;
    J.cond = cond;
    J.stats = stats;
    J.next = 0;
    J.bad = 0;
    pthread_mutex_init(&J.mu, NULL);
    if (threads <= 0) threads = (int)sysconf(_SC_NPROCESSORS_ONLN);
    if (threads < 1) threads = 1;
    if ((size_t)threads > n_cw) threads = (int)(n_cw ? n_cw : 1);
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)threads);
    for (int t = 1; t < threads; t++) pthread_create(&th[t], NULL, og_worker, &J);
    og_worker(&J);
    for (int t = 1; t < threads; t++) pthread_join(th[t], NULL);
    free(th);
    pthread_mutex_destroy(&J.mu);
    return J.bad ? -1 : 0;
}

int og_decode_cubes_cond(const double *llr, size_t n_cw, int L, int Imax, uint64_t Tmax,
                         double p_et, double alpha, uint8_t *bits, double *soft, double *cond,
                         og_cw_stats *stats, int threads)
{
    return og_decode_product(3, llr, n_cw, L, Imax, Tmax, p_et, alpha, bits, soft, cond, stats,
                             threads);
}

int og_decode_squares(const double *llr, size_t n_cw, int L, int Imax, uint64_t Tmax,
                      double p_et, double alpha, uint8_t *bits, double *soft, og_cw_stats *stats,
                      int threads)
{
    return og_decode_product(2, llr, n_cw, L, Imax, Tmax, p_et, alpha, bits, soft, NULL, stats,
                             threads);
}
