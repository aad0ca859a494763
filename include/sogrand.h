/*
 * sogrand.h - C ABI of the MI355X (gfx950) SOGRAND decoder for the (16,8)^3
 * cubic product code.
 *
 * All entry points are extern "C", take plain pointers and sizes, and return
 * 0 on success or a negative SOGRAND_E* code (never exit()). Buffers may be
 * host memory (pageable or pinned) or device memory; the library tells them
 * apart with hipPointerGetAttributes. Results never depend on how codewords
 * are batched or which backend decodes them.
 *
 * Data layout (the reference decoder's file formats):
 *   LLR cube  : 4096 float64 per codeword, index a0 + 16*a1 + 256*a2
 *               (SOGRAND_C/cubic_decoder_final.c:119, cubic_encoder.c:89-97)
 *   info bits : 512 per codeword, bit b = a0 + 8*a1 + 64*a2 (a0,a1,a2 < 8),
 *               packed MSB-first, 64 bytes per codeword (cubic_decoder_final.c:267-280)
 */
#ifndef SOGRAND_H
#define SOGRAND_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SOGRAND_N 16
#define SOGRAND_K 8
#define SOGRAND_CUBE 4096     /* LLRs per codeword          */
#define SOGRAND_INFO_BITS 512 /* info bits per codeword     */
#define SOGRAND_INFO_BYTES 64 /* packed info bytes per cw   */
#define SOGRAND_MAX_L 3       /* largest list size supported (the reference uses 3) */

/* error codes */
#define SOGRAND_OK 0
#define SOGRAND_EINVAL (-1)     /* bad argument                               */
#define SOGRAND_ENOMEM (-2)     /* host or device allocation failed          */
#define SOGRAND_EHIP (-3)       /* a HIP runtime call failed                   */
#define SOGRAND_ENONFINITE (-4) /* not returned: every double is decoded as the
                                   reference decodes it (+-inf and NaN included) */
#define SOGRAND_ENODEV (-5)     /* no usable gfx950 device                     */
#define SOGRAND_EIO (-6)        /* file open/read/write failed (CLI helpers)   */
#define SOGRAND_EBUSY (-7)      /* serving queue full / admission refused      */

/* precision */
#define SOGRAND_FP64 0       /* bit-exact with SOGRAND_C                              */
#define SOGRAND_FP32_MIXED 1 /* FP32 sort + pattern metrics, FP64 list normalisation */

/* decoder variant: which of the reference's SOGRAND_C decoders to reproduce */
#define SOGRAND_VARIANT_FINAL 0  /* cubic_decoder_final.c (with `epsilon` = DBL_EPSILON:
                                    cubic_decoder_updated.c)                          */
#define SOGRAND_VARIANT_LEGACY 1 /* cubic_decoder.c, the earlier decoder: JacLog cut at
                                    +-30, IC guard, P_notGuess never lowered (every
                                    search fills its list), clamps 1e-9 and 1e-30;
                                    FP64 only, `epsilon` unused                       */

/* backend */
#define SOGRAND_BACKEND_PLAIN 0      /* one workgroup per codeword, static lines */
#define SOGRAND_BACKEND_PERSISTENT 1 /* persistent waves + device queue          */
#define SOGRAND_BACKEND_AUTO 3       /* by batch size: plain up to one cube per resident
                                        persistent wave, then persistent with one cube
                                        slot per wave, two from twice that on */
#define SOGRAND_BACKEND_LANEGROUP 2  /* one 1024-lane workgroup per codeword, the
                                        cube in LDS, four lanes per line: the
                                        shortest time per codeword (small frames);
                                        FP64 with the final variant only         */

/* Decoder parameters. Defaults (sogrand_params_default) are the reference's
 * constants: L=3, Imax=30, Tmax=UINT64_MAX, p_ET=1e-5, alpha=0.7, EPSILON=1e-40
 * (SOGRAND_C/cubic_decoder_final.c:16, 94-102). */
typedef struct sogrand_params {
    uint32_t L;         /* list size, 1..SOGRAND_MAX_L                                  */
    uint32_t Imax;      /* maximum product-code iterations (3 phases each), 1..1000     */
    uint64_t Tmax;      /* guess budget per component decode, checked once per weight
                           level as the reference does; 0 means 0x7fffffff              */
    double p_ET;        /* list confidence threshold is 1 - p_ET                         */
    double alpha;       /* extrinsic scaling                                             */
    int32_t precision;  /* SOGRAND_FP64 or SOGRAND_FP32_MIXED                            */
    int32_t backend;    /* SOGRAND_BACKEND_*                                             */
    int32_t wave_slots; /* persistent: cube slots per wave, 1..4 (0 = tuned default, 2)   */
    int32_t waves_per_cu; /* persistent: resident waves per CU (0 = as many as fit)      */
    int32_t lane_batch; /* persistent: free lanes gathered before new lines are handed out
                           (0 = tuned default)                                            */
    int32_t step_len;   /* persistent: guesses per lane between scheduling points, taken
                           in steps of 16 (0 = tuned default)                             */
    int32_t use_graph;  /* 1: when every buffer of a decode is device memory, capture its
                           launch sequence in a hipGraph once per (buffers, size,
                           parameters) and replay it on later calls                      */
    double epsilon;     /* EPSILON of the list confidence test and the soft-output clamps:
                           1e-40 as cubic_decoder_final.c:16 (default); DBL_EPSILON gives
                           the reference's cubic_decoder_updated.c:16 variant            */
    int32_t variant;    /* SOGRAND_VARIANT_* (default SOGRAND_VARIANT_FINAL)             */
} sogrand_params;

/* Per-codeword statistics (the reference prints only their averages,
 * cubic_decoder_final.c:299-304). */
typedef struct sogrand_cw_stats {
    uint32_t half_iters; /* phases run; the reference's n_iter = half_iters / 2 */
    uint32_t converged;  /* 1 if the early-termination test passed              */
    uint64_t NG;         /* guesses summed over all component decodes           */
    uint64_t NG_p;       /* per-phase maximum guesses, summed over phases       */
} sogrand_cw_stats;

typedef struct sogrand_ctx sogrand_ctx;

void sogrand_params_default(sogrand_params *p);
const char *sogrand_strerror(int code);

/* Create a decoder context on HIP device `device`. Uploads the error-pattern
 * schedule tables. Replaces the reference's per-process set-up
 * (cubic_decoder1.cu:950-985: getGH_sys_CRC, cudaMemcpyToSymbol). */
int sogrand_ctx_create(sogrand_ctx **out, int device, const sogrand_params *params);
int sogrand_ctx_set_params(sogrand_ctx *ctx, const sogrand_params *params);
void sogrand_ctx_destroy(sogrand_ctx *ctx);

/* Decode n_cw cubes. llr: 4096 f64 per codeword. info_bits: 64 B per codeword
 * (required). soft_out (optional, may be NULL): final L_APP, 4096 f64 per
 * codeword. stats (optional): one per codeword. hip_stream: hipStream_t or NULL
 * for the context's stream. Synchronous: returns when results are written.
 * Replaces the reference's decode loop, cubic_decoder_final.c:119-289, and the
 * GPU launch + copy-back, cubic_decoder1.cu:1059-1110 (kernel boundary
 * highly_optimized_sogrand_kernel, cubic_decoder1.cu:635-643). */
int sogrand_decode_cubic(sogrand_ctx *ctx, const double *llr, size_t n_cw, uint8_t *info_bits,
                         double *soft_out, sogrand_cw_stats *stats, void *hip_stream);

/* sogrand_decode_cubic without any allocation. It has no counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * llr, info_bits and the optional soft_out / stats must be device memory
 * (SOGRAND_EINVAL otherwise). It runs the plain kernel whatever the context's
 * backend (results are the same bit for bit), with at most
 * SOGRAND_NOALLOC_SCRATCH_BYTES / 32 KiB = 1024 workgroups, each looping over
 * codewords with one 32 KiB scratch cube; the scratch is a static array of
 * SOGRAND_NOALLOC_SCRATCH_BYTES that every GPU the library runs on holds until
 * the process ends, so it fits any n_cw. All calls on one GPU share that array:
 * calls that overlap on one GPU (from several contexts or threads) race.
 * Synchronous, like sogrand_decode_cubic. */
#define SOGRAND_NOALLOC_SCRATCH_BYTES (32u << 20)
int sogrand_decode_cubic_noalloc(sogrand_ctx *ctx, const double *llr, size_t n_cw,
                                 uint8_t *info_bits, double *soft_out, sogrand_cw_stats *stats,
                                 void *hip_stream);

/* Decode n_cw squares of the (16,8)^2 product code: 256 f64 LLRs per square
 * (index a0 + 16*a1, the layout SOGRAND_C/channel_sim.c writes for dimension
 * 2), 8 info bytes out per square (info bits b = a0 + 8*a1, MSB first, as
 * cubic_decoder_final.c:267-280 packs the cube's), optional final L_APP (256
 * f64 per square) and statistics, with the context's parameters. The square
 * decode is the reference's per-codeword loop (cubic_decoder_final.c:119-289)
 * over two axes: phases alternate rows and columns, half an iteration each, and
 * stop when every row and column of the hard decision is a component codeword.
 * No counterpart in the reference: its channel simulator accepts the square
 * code (SOGRAND_C/channel_sim.c:26-29) but it ships no square decoder. Host or
 * device pointers; synchronous; the backend parameter does not apply (one
 * kernel: sixteen squares per workgroup). */
int sogrand_decode_square(sogrand_ctx *ctx, const double *llr, size_t n_cw, uint8_t *info_bits,
                          double *soft_out, sogrand_cw_stats *stats, void *hip_stream);

/* Component SISO decode of n_lines independent 16-value lines on the GPU:
 * L_APP, L_E (16 f64 per line) and N_guess per line. Replaces SOGRAND_bitSO
 * (cubic_decoder_final.c:521) with L, Tmax, thres from the context.
 * L_APP/L_E may be NULL. */
int sogrand_bitso_lines(sogrand_ctx *ctx, const double *llr, size_t n_lines, double *L_APP,
                        double *L_E, uint32_t *N_guess, void *hip_stream);

/* Time (ms, HIP events) of the decode kernels of the last sogrand_decode_cubic
 * call on this context, from the first kernel's start to the last kernel's end
 * (with host buffers, chunks are decoded on two alternating streams and their
 * launches overlap), and how many launches it made. */
int sogrand_last_kernel_ms(sogrand_ctx *ctx, double *ms, uint32_t *launches);

/* With use_graph: how many times the context's current captured graph has been
 * replayed since it was captured (0 right after a capture). */
int sogrand_graph_replays(sogrand_ctx *ctx, uint32_t *replays);

/* The error-pattern schedule the decoder streams for intercept ic (ic >= 55
 * all share one order): 65535 16-bit masks over sorted positions (bit b = b-th
 * least reliable bit) in the order the reference's landslide loop visits them
 * (cubic_decoder_final.c:642-744, mountain_build :775-789), then a 0. out holds
 * 65536 entries. Host only. */
int sogrand_schedule_table(int32_t ic, uint16_t *out);

/* ---- synthetic data, restating the reference's channel (channel_sim.c:34-39,61-89)
 * with a seeded counter-based generator. Codeword c of a stream depends only on
 * (seed, c): any chunking gives identical data; host and device agree bit for bit. */

/* sigma of BPSK over AWGN at Eb/N0 (dB) for the rate-1/8 cube code
 * (channel_sim.c:34-39: EsN0 = EbN0 + 10 log10(2R), sigma = 1/sqrt(10^(EsN0/10))). */
double sogrand_sigma_for_ebn0(double ebn0_db);

/* Info bits of codewords [cw_offset, cw_offset + n_cw): 64 B per codeword. */
int sogrand_gen_info_bits(uint64_t seed, uint64_t cw_offset, size_t n_cw, uint8_t *info_bits);

/* Systematic (16,8)^3 encoding of packed info bits into the reference
 * encoder's output format: 4096 code bits per codeword, index a0+16*a1+256*a2,
 * packed MSB-first, 512 B per codeword (cubic_encoder.c:23-100). Host only. */
int sogrand_encode_cubes(const uint8_t *info_bits, size_t n_cw, uint8_t *code_bytes);

/* LLRs (4096 f64 per codeword) for codewords [cw_offset, cw_offset + n_cw):
 * info bits from (seed, c), encoded, BPSK 0 -> +1, AWGN with sigma[c - cw_offset]
 * (n_sigma == n_cw) or sigma[0] (n_sigma == 1), LLR = 2y / sigma^2.
 * sogrand_gen_llr runs on the host; sogrand_gen_llr_device writes device
 * memory on the GPU (llr, sigma device or host pointers; hip_stream or NULL). */
int sogrand_gen_llr(uint64_t seed, uint64_t cw_offset, size_t n_cw, const double *sigma,
                    size_t n_sigma, double *llr);
int sogrand_gen_llr_device(sogrand_ctx *ctx, uint64_t seed, uint64_t cw_offset, size_t n_cw,
                           const double *sigma, size_t n_sigma, double *llr, void *hip_stream);

/* ---- serving loop: continuous batching of decode requests ----
 * No counterpart in the reference, whose nearest relative is the multi-stream
 * CUDA decoder (SOGRAND_CUDA/cubic_decoder.cu:1145-1179, batches written out of
 * order). A worker thread keeps up to max_inflight batches decoding at once,
 * each on its own stream with a fixed share of hbm_budget (no allocation after
 * create). Whenever a batch slot frees up, every waiting request that fits is
 * gathered into the next batch, so requests that arrive while others decode are
 * admitted as slots free up and their work overlaps on the device. Results are
 * reported by request id; each codeword decodes exactly as it does alone. The
 * server copies the context's parameters at create; the context must outlive it. */
typedef struct sogrand_server sogrand_server;

typedef struct sogrand_server_cfg {
    size_t hbm_budget;       /* device bytes for all batch slots (default 4 GiB)          */
    uint32_t max_inflight;   /* batch slots = concurrent launches, 1..32 (default 16; they
                                run concurrently up to the HIP runtime's hardware queues,
                                GPU_MAX_HW_QUEUES)                                          */
    uint32_t max_batch_cw;   /* codewords per batch (0 = what the budget allows, <= 65536) */
    uint32_t max_pending_cw; /* waiting codewords before submit refuses (default 2^20)     */
    int32_t with_soft;       /* reserve room for soft outputs (default 0)                  */
    uint32_t plain_below_cw; /* batches with fewer codewords run the plain kernel, which
                                finishes a cube sooner when the device has room
                                (0 = 3 per CU; 1 = always the configured backend)         */
    uint32_t hold_above_cw;  /* no batch is launched while this many codewords are
                                decoding, so that waiting requests leave together as one
                                larger batch (0 = no limit). A batch completes with its
                                slowest cube, so holding costs throughput on uneven work  */
    uint32_t lanegroup_max_cw; /* a batch runs the lane-group kernel (cube in LDS, four
                                lanes per line: a codeword finishes soonest) while it and
                                every batch still decoding hold at most this many
                                codewords; FP64 final variant only (default 0 = never:
                                at 70 % load its CU-sized workgroups lengthen the tail)    */
    uint32_t small_batch_cw; /* while at most backlog_cw codewords wait, a batch takes at
                                most this many (at least one request), so that a frame
                                completes with its own slowest codeword rather than the
                                slowest of a large batch (default 256; 0 = off)           */
    uint32_t backlog_cw;     /* waiting codewords above which a batch takes everything that
                                fits, for throughput (0 = 8 x small_batch_cw)             */
} sogrand_server_cfg;

typedef struct sogrand_completion {
    uint64_t id;     /* the id given to sogrand_server_submit                         */
    int32_t status;  /* SOGRAND_OK or SOGRAND_EHIP                                      */
    uint32_t n_cw;
    double t_submit; /* seconds on the server clock: submitted, batch launched, results */
    double t_start;  /* copied back                                                     */
    double t_done;
} sogrand_completion;

void sogrand_server_cfg_default(sogrand_server_cfg *cfg);
int sogrand_server_create(sogrand_server **out, sogrand_ctx *ctx, const sogrand_server_cfg *cfg);
/* Codewords per batch slot (the largest request the server accepts). */
int sogrand_server_batch_cw(const sogrand_server *srv, size_t *batch_cw);
/* Queue a request of n_cw codewords (llr: host or device memory, 4096 f64 each).
 * Outputs (host or device) are written before its completion is reported; the
 * input must stay valid until then. soft_out needs with_soft. Returns
 * SOGRAND_EBUSY when max_pending_cw codewords are already waiting. */
int sogrand_server_submit(sogrand_server *srv, uint64_t id, const double *llr, size_t n_cw,
                          uint8_t *info_bits, double *soft_out, sogrand_cw_stats *stats);
/* Move up to max completions into out (oldest first); waits up to wait_ms for
 * the first one when none is ready. */
int sogrand_server_poll(sogrand_server *srv, sogrand_completion *out, size_t max, size_t *n_out,
                        int wait_ms);
/* Block until every submitted request has completed. */
int sogrand_server_drain(sogrand_server *srv);
double sogrand_server_clock(const sogrand_server *srv);
/* Finishes the queued work, then frees the server. */
void sogrand_server_destroy(sogrand_server *srv);

#ifdef __cplusplus
}
#endif
#endif
