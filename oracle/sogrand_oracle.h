/*
 * CPU oracle for the SOGRAND (16,8)^3 cubic product-code decoder.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C restatement of the reference CPU
 * decoder SOGRAND_C/cubic_decoder_final.c, used by tests/, the smoke check and
 * the CPU-baseline leg of bench.py as the thing the HIP path is checked against.
 * The product (libsogrand) never links, loads or calls it.
 */
#ifndef SOGRAND_ORACLE_H
#define SOGRAND_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OG_N 16
#define OG_K 8
#define OG_CUBE 4096
#define OG_INFO 512
#define OG_MAXL 8

/* Trace of one component decode (for line-level golden comparisons). */
typedef struct {
    int32_t ic;              /* intercept used for the schedule                   */
    int32_t n_list;          /* number of list entries (0..L)                      */
    uint32_t T;              /* guesses made, including the hard decision          */
    uint32_t pad0;
    double pnl;              /* P_neg of the last list insertion (0 if none)       */
    double pm[OG_MAXL];      /* path metric of each list entry                     */
    double conf[OG_MAXL];    /* confidence returned at each insertion              */
    uint32_t t_hit[OG_MAXL]; /* value of T when each entry was found               */
    int32_t wt_hit[OG_MAXL]; /* logistic weight of each entry (0 for the HD guess) */
    uint16_t cw[OG_MAXL];    /* codeword, natural order, bit i = position i        */
    uint8_t perm[OG_N];      /* sorted position -> natural position                */
} og_line_trace;

typedef struct {
    uint32_t half_iters; /* phases run (the reference prints half of this as n_iter) */
    uint32_t converged;  /* 1 if the early-termination test passed                    */
    uint64_t NG;         /* total guesses over all component decodes                  */
    uint64_t NG_p;       /* sum over phases of the maximum per-line guess count       */
} og_cw_stats;

/* Code tables of the (16,8) systematic CRC code (poly 0xeb). G, H row-major 8x16;
 * hcol[p] = column p of H as an 8-bit syndrome (bit j = row j). */
void og_code_tables(uint8_t *G, uint8_t *H, uint8_t *hcol);

/* Ascending in-place sort of a[0..n) carrying perm, with the reference's tie order. */
void og_sort_with_perm(double *a, uint8_t *perm, int n);

/* First `count` entries of the error-pattern schedule for intercept ic, as
 * 16-bit masks over sorted positions (bit b = b-th least reliable bit). Returns
 * how many were written (the schedule holds 65535 nonzero patterns). */
int og_schedule_prefix(int32_t ic, uint16_t *masks, int count);

/* EPSILON of the soft output and the confidence test (default 1e-40, the value
 * of cubic_decoder_final.c:16; DBL_EPSILON is cubic_decoder_updated.c's). Applies
 * to every later decode in the process. */
void og_set_epsilon(double eps);

/* Decoder variant for every later decode in the process: 0 = the semantics of
 * cubic_decoder_final.c (and of _updated.c with its EPSILON), 1 = the
 * reference's earlier cubic_decoder.c (JacLog cut at +-30, IC guard
 * beta > 1e-9, P_notGuess never decremented for this code so the search runs
 * until the list is full, P_neg of the last guess's weight when no hit ends
 * the search, clamps 1e-9 and 1e-30). */
void og_set_variant(int variant);

/* Component SISO decode of one line (any doubles, +-inf and NaN as the
 * reference takes them). Returns 0.
 * L_APP / L_E may be NULL; trace may be NULL. */
int og_siso(const double *llr, int L, uint64_t Tmax, double thres, double *L_APP, double *L_E,
            og_line_trace *trace);

/* 1 if the hard decision of the cube (bit = !(x > 0)) is a product codeword. */
int og_cube_is_codeword(const double *L_APP);

/* The same for a square of the (16,8)^2 code (256 f64, index a0 + 16*a1). */
int og_square_is_codeword(const double *L_APP);

/* Decode n_cw cubes (4096 f64 each, index a0 + 16*a1 + 256*a2).
 * bits: 64 B per codeword, MSB-first packed info bits b = a0 + 8*a1 + 64*a2.
 * soft (optional): final L_APP, 4096 f64 per codeword. stats (optional).
 * threads <= 0 means all online CPUs. Returns 0. */
int og_decode_cubes(const double *llr, size_t n_cw, int L, int Imax, uint64_t Tmax, double p_et,
                    double alpha, uint8_t *bits, double *soft, og_cw_stats *stats, int threads);

/* As og_decode_cubes, plus cond (optional, 4096 per codeword): for each final
 * L_APP value, the relative rounding-error bound of P_notGuess in the line
 * decode that produced it (the reference's own arithmetic determines that
 * L_APP only to about this many units). */
int og_decode_cubes_cond(const double *llr, size_t n_cw, int L, int Imax, uint64_t Tmax,
                         double p_et, double alpha, uint8_t *bits, double *soft, double *cond,
                         og_cw_stats *stats, int threads);

/* Decode n_cw squares of the (16,8)^2 product code (256 f64 each, index
 * a0 + 16*a1) with the same per-codeword loop over two axes: phases alternate
 * rows (axis 0) and columns (axis 1), n_iter counts half an iteration per
 * phase, early termination when every row and column is a component codeword.
 * bits: 8 B per codeword, MSB-first info bits b = a0 + 8*a1. soft (optional):
 * final L_APP, 256 f64 per codeword. stats (optional). Returns 0. The
 * reference ships no square decoder (its channel simulator accepts dimension
 * 2, channel_sim.c:26-29); this is the cubic loop restricted to two axes. */
int og_decode_squares(const double *llr, size_t n_cw, int L, int Imax, uint64_t Tmax,
                      double p_et, double alpha, uint8_t *bits, double *soft, og_cw_stats *stats,
                      int threads);

#ifdef __cplusplus
}
#endif
#endif
