/*
 * Harness around the reference decoder's own functions (test infrastructure).
 *
 * Linked only against SOGRAND_C/cubic_decoder_final.c compiled unmodified with
 * its main() renamed (see Makefile). Nothing here re-implements the decoder: every
 * entry point forwards to a reference function so that Python tests can record
 * golden vectors from the reference itself.
 *
 *   refh_code_tables   -> getGH_sys_CRC          (cubic_decoder_final.c:468)
 *   refh_siso          -> SOGRAND_bitSO          (cubic_decoder_final.c:521)
 *   refh_main_logic    -> sogrand_main_logic     (cubic_decoder_final.c:587)
 *   refh_quicksort     -> QuickSort              (cubic_decoder_final.c:763)
 *   refh_early_term    -> early_termination      (cubic_decoder_final.c:317)
 *   refh_decode_cube   -> the per-codeword loop of main (cubic_decoder_final.c:119-289),
 *                         driving the reference SOGRAND_bitSO / early_termination
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* Layout of the reference's Tensor_d (cubic_decoder_final.c:20-23), needed to
 * pass a cube to early_termination by value. */
typedef struct {
    double *data;
    int dim1, dim2, dim3;
} refh_tensor_d;

void getGH_sys_CRC(int n, int k, int **G, int **H);
void SOGRAND_bitSO(double *L_APP, double *L_E, int *N_guess, double *llr, int **H_matrix,
                   int n, int k, int L, uint64_t Tmax, double thres, int even);
void sogrand_main_logic(double *chat_list, double *s_list, double *T_val, double *curL_val,
                        double *pNL_val, double *APP_list, double *llr, uint8_t *H_flat, int n,
                        int s, int IC, uint64_t L, uint64_t Tmax, double thres, int even);
void QuickSort(double *a, size_t *perm, uint64_t n);
int early_termination(refh_tensor_d L_APP, int **G, int n, int k);

static int refh_G_store[8][16], refh_H_store[8][16];
static int *refh_G[8], *refh_H[8];
static int refh_ready = 0;

static void refh_init(void)
{
    if (refh_ready) return;
    for (int i = 0; i < 8; i++) {
        refh_G[i] = refh_G_store[i];
        refh_H[i] = refh_H_store[i];
    }
    getGH_sys_CRC(16, 8, refh_G, refh_H);
    refh_ready = 1;
}

/* G (8x16) and H (8x16) as produced by the reference, row-major ints. */
void refh_code_tables(int *G_out, int *H_out)
{
    refh_init();
    for (int i = 0; i < 8; i++)
        for (int j = 0; j < 16; j++) {
            G_out[i * 16 + j] = refh_G[i][j];
            H_out[i * 16 + j] = refh_H[i][j];
        }
}

/* One component decode exactly as main() calls it (L=3, even=0). */
void refh_siso(const double *llr, double *L_APP, double *L_E, int *N_guess, uint64_t Tmax,
               double thres)
{
    refh_init();
    double in[16];
    memcpy(in, llr, sizeof in);
    SOGRAND_bitSO(L_APP, L_E, N_guess, in, refh_H, 16, 8, 3, Tmax, thres, 0);
}

/* Raw list search with caller-chosen parity checks: H_flat is s x n (u8). */
void refh_main_logic(const double *llr, const uint8_t *H_flat, int s, int IC, uint64_t L,
                     uint64_t Tmax, double thres, double *chat_list /* n*L */,
                     double *s_list /* 4*L */, double *T, double *curL, double *pNL)
{
    double in[16];
    memcpy(in, llr, sizeof in);
    uint8_t *h = (uint8_t *)malloc((size_t)s * 16);
    memcpy(h, H_flat, (size_t)s * 16);
    double *app = (double *)calloc(L ? L : 1, sizeof(double));
    *curL = 0.0;
    sogrand_main_logic(chat_list, s_list, T, curL, pNL, app, in, h, 16, s, IC, L, Tmax, thres, 0);
    free(app);
    free(h);
}

void refh_quicksort(double *a, uint64_t *perm, uint64_t n)
{
    size_t p[64];
    for (uint64_t i = 0; i < n; i++) p[i] = (size_t)perm[i];
    QuickSort(a, p, n);
    for (uint64_t i = 0; i < n; i++) perm[i] = (uint64_t)p[i];
}

int refh_early_term(const double *L_APP)
{
    refh_init();
    refh_tensor_d t;
    t.data = (double *)malloc(4096 * sizeof(double));
    memcpy(t.data, L_APP, 4096 * sizeof(double));
    t.dim1 = t.dim2 = t.dim3 = 16;
    int r = early_termination(t, refh_G, 16, 8);
    free(t.data);
    return r;
}

/* Per-codeword decode driving the reference component decoder and stopping
 * rule. Outputs: final L_APP (4096), half-iterations, NG, NG_p, 512 info bits
 * (one per byte, b = a0 + 8*a1 + 64*a2). Cube index = a0 + 16*a1 + 256*a2. */
void refh_decode_cube(const double *L_ch, int Imax, uint64_t Tmax, double alpha,
                      double *L_APP, int *half_iters, int64_t *NG, int64_t *NG_p,
                      uint8_t *bits)
{
    refh_init();
    const double thres = 1.0 - 1e-5;
    double *L_E = (double *)calloc(4096, sizeof(double));
    double *inp = (double *)malloc(4096 * sizeof(double));
    int64_t ng = 0, ngp = 0;
    int halves = 0;
    static const int stride[3] = {1, 16, 256};
    static const int o1[3] = {16, 1, 1};
    static const int o2[3] = {256, 256, 16};
    memcpy(L_APP, L_ch, 4096 * sizeof(double));
    for (int it = 1; it <= Imax; it++) {
        int done = 0;
        for (int ax = 0; ax < 3; ax++) {
            halves++;
            for (int i = 0; i < 4096; i++) {
                double t = alpha * L_E[i];
                inp[i] = L_ch[i] + t;
            }
            int ngmax = 0;
            for (int y = 0; y < 16; y++)
                for (int x = 0; x < 16; x++) {
                    int base = x * o1[ax] + y * o2[ax];
                    double v[16], la[16], le[16];
                    for (int j = 0; j < 16; j++) v[j] = inp[base + j * stride[ax]];
                    int ngl = 0;
                    SOGRAND_bitSO(la, le, &ngl, v, refh_H, 16, 8, 3, Tmax, thres, 0);
                    ng += ngl;
                    if (ngl > ngmax) ngmax = ngl;
                    for (int j = 0; j < 16; j++) {
                        L_APP[base + j * stride[ax]] = la[j];
                        L_E[base + j * stride[ax]] = le[j];
                    }
                }
            ngp += ngmax;
            refh_tensor_d t = {L_APP, 16, 16, 16};
            if (early_termination(t, refh_G, 16, 8)) {
                done = 1;
                break;
            }
        }
        if (done) break;
    }
    for (int a2 = 0; a2 < 8; a2++)
        for (int a1 = 0; a1 < 8; a1++)
            for (int a0 = 0; a0 < 8; a0++)
                bits[a0 + 8 * a1 + 64 * a2] = (L_APP[a0 + 16 * a1 + 256 * a2] > 0.0) ? 0 : 1;
    *half_iters = halves;
    *NG = ng;
    *NG_p = ngp;
    free(L_E);
    free(inp);
}
