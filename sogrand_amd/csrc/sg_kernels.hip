// HIP kernels of the SOGRAND (16,8)^3 decoder for gfx950.
//
// plain_kernel   : one 256-lane workgroup per cube; lane t owns line t of every
//                  phase; L_E in a per-workgroup global slot, hard decisions in
//                  LDS; phases separated by workgroup barriers. Replaces the
//                  reference GPU kernel highly_optimized_sogrand_kernel
//                  (SOGRAND_CUDA/cubic_decoder1.cu:635-841) with the CPU
//                  decoder's semantics (SOGRAND_C/cubic_decoder_final.c:119-289).
// square_kernel  : the (16,8)^2 square code: sixteen squares per workgroup, a
//                  16-lane group per square, one lane per line, phases
//                  alternating rows and columns.
// lines_kernel   : SOGRAND_bitSO (cubic_decoder_final.c:521) on independent lines.
// gen_kernel     : seeded info bits -> product encoding -> BPSK/AWGN -> LLR.
#include <hip/hip_runtime.h>

#include "sg_common.h"
#include "sg_device.h"
#include "sg_launch.h"

#pragma clang fp contract(off)

namespace sg {

struct CwStatsDev {
    uint32_t half_iters;
    uint32_t converged;
    uint64_t NG;
    uint64_t NG_p;
};

// Axis geometry: line (x, y) of axis ax covers base + j*stride, j = 0..15.
// axis 0: (a1, a2) over a0; axis 1: (a0, a2) over a1; axis 2: (a0, a1) over a2.
__device__ __forceinline__ void axis_line(int ax, int line, int &base, int &stride)
{
    const int x = line & 15, y = line >> 4;
    if (ax == 0) {
        base = 16 * x + 256 * y;
        stride = 1;
    } else if (ax == 1) {
        base = x + 256 * y;
        stride = 16;
    } else {
        base = x + 16 * y;
        stride = 256;
    }
}

// Component input of position p: L_channel + alpha * L_E, a multiply then an
// add (:147-155). In the first phase L_E is all zeros (:129).
struct CubeSrc {
    const double *lch;
    const double *le;
    double alpha;
    int base, stride;
    bool first;
    __device__ double operator()(int i) const
    {
        const int p = base + i * stride;
        const double e = first ? 0.0 : le[p];
        const double t = alpha * e;
        return lch[p] + t;
    }
};

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        uint32_t w = __shfl_xor(v, o, 64);
        v = w > v ? w : v;
    }
    return v;
}

constexpr int kPlainThreads = 256;

template <bool FINITE_T, class P>
__global__ __launch_bounds__(kPlainThreads) void plain_kernel(CubeJob job)
{
    using Rec = typename P::Rec;
    __shared__ Rec s_delta[17 * kPlainThreads];
    __shared__ uint16_t s_hp[17 * kPlainThreads];
    __shared__ uint8_t s_hd[kCube];
    __shared__ uint32_t s_sum[kPlainThreads / 64], s_max[kPlainThreads / 64];
    __shared__ ExpEntry s_exp[kLdsTableEntries];

    const int t = threadIdx.x;
    load_exp_table(s_exp, t, job.eps, job.thres);
    __syncthreads();
    const LineLdsT<Rec, kPlainThreads> S(s_delta, s_hp, t);
    const DecodeCfg cfg{job.tmax, job.thres, job.sched, job.L, s_exp, kLogTable, job.eps, job.eps_log};
    double *le = job.le_scratch + (size_t)blockIdx.x * kCube;

    for (size_t cube = blockIdx.x; cube < job.n_cw; cube += gridDim.x) {
        const double *lch = job.llr + cube * kCube;
        double *soft = job.soft ? job.soft + cube * kCube : nullptr;
        uint32_t half = 0, conv = 0;
        uint64_t ng = 0, ngp = 0;
        for (uint32_t it = 1; it <= job.Imax && !conv; ++it) {
            for (int ax = 0; ax < 3; ++ax) {
                ++half; // n_iter += 0.5 before each phase (:145, :184, :223)
                CubeSrc src{lch, le, job.alpha, 0, 0, it == 1 && ax == 0};
                axis_line(ax, t, src.base, src.stride);
                uint32_t hd, s0;
                double pm_hd;
                int32_t ic;
                LineList ls;
                line_setup_p<P>(src, S, hd, s0, pm_hd, ic, cfg);
                line_search<FINITE_T, P>(S, hd, s0, pm_hd, ic, cfg, ls);
                const double pnotin = fmax(ls.pnl, cfg.eps);
                ListEntries ent;
                list_entries<P>(ls, S, hd, pm_hd, ic, cfg, ent);
#pragma unroll 1
                for (int i = 0; i < 16; ++i) {
                    const int p = src.base + i * src.stride;
                    const double xin = src(i);
                    double la, lenew;
                    if (ls.nl == 0) { // empty list: L_APP = llr, L_E = 0 (:537-539)
                        la = xin;
                        lenew = 0.0;
                    } else {
                        la = soft_app(ent, ls.nl, pnotin, i, xin, cfg);
                        lenew = la - xin;
                    }
                    le[p] = lenew;
                    if (soft) soft[p] = la;
                    s_hd[p] = (la > 0.0) ? 0 : 1;
                }
                // NG += N_guess; NG_p += max over the phase's lines (:162-163, :170)
                const uint32_t ws = wave_sum_u32(ls.T), wm = wave_max_u32(ls.T);
                if ((t & 63) == 0) {
                    s_sum[t >> 6] = ws;
                    s_max[t >> 6] = wm;
                }
                __syncthreads();
                uint32_t bs = 0, bm = 0;
#pragma unroll
                for (int w = 0; w < kPlainThreads / 64; ++w) {
                    bs += s_sum[w];
                    bm = s_max[w] > bm ? s_max[w] : bm;
                }
                ng += bs;
                ngp += bm;
                // early termination: the hard decision is a product codeword iff
                // all 768 lines are component codewords (equivalent to the
                // re-encode test of :317-391)
                bool line_bad = false;
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    int b0, st;
                    axis_line(a, t, b0, st);
                    uint32_t w = 0;
#pragma unroll
                    for (int j = 0; j < 16; ++j) w |= (uint32_t)s_hd[b0 + j * st] << j;
                    line_bad = line_bad || (syndrome16(w) != 0);
                }
                const int any_bad = __syncthreads_or(line_bad);
                if (!any_bad) {
                    conv = 1;
                    break;
                }
            }
        }
        if (t < kInfoBytes) { // info bit b = a0 + 8*a1 + 64*a2, MSB first (:267-280)
            uint32_t byte = 0;
            const int a1 = t & 7, a2 = t >> 3;
#pragma unroll
            for (int q = 0; q < 8; ++q) byte |= (uint32_t)s_hd[q + 16 * a1 + 256 * a2] << (7 - q);
            job.info[cube * kInfoBytes + t] = (uint8_t)byte;
        }
        if (t == 0) {
            if (job.stats) {
                CwStatsDev *st = reinterpret_cast<CwStatsDev *>(job.stats) + cube;
                st->half_iters = half;
                st->converged = conv;
                st->NG = ng;
                st->NG_p = ngp;
            }
        }
        __syncthreads(); // s_hd is read above before the next cube overwrites it
    }
}

// The (16,8)^2 square code (256 positions a0 + 16 a1, info bits b = a0 + 8 a1):
// the reference's per-codeword loop (cubic_decoder_final.c:119-289) over two
// axes. The reference ships no square decoder; its channel simulator accepts
// the square code (SOGRAND_C/channel_sim.c:26-29). Phases alternate rows
// (axis 0: line i = positions 16 i + j) and columns (axis 1: i + 16 j), half an
// iteration each, with early termination when every row and column of the
// hard decision is a component codeword. A workgroup takes sixteen squares at
// a time, the 16-lane group 16 g .. 16 g + 15 square g (groups never straddle
// a wave, so a group's lines exchange hard decisions by ballots, and groups run
// independently: no workgroup barrier after the table load). L_E sits in the
// workgroup's 4096-double scratch slot, 256 doubles per square; a workgroup
// fence orders a phase's stores before the next phase's loads, which other
// lanes of the same wave make.
constexpr int kSquareThreads = 256;
constexpr int kSquarePos = 256;
constexpr int kSquareInfoBytes = 8;

// For the 16-lane group of the calling lane: bit j of the result, in lane
// 16 g + i, is bit i of lane 16 g + j's word w (the hard decisions of the
// group's lines transposed: rows <-> columns).
__device__ __forceinline__ uint32_t group16_transpose(uint32_t w, int i, int gshift)
{
    uint32_t out = 0;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const uint32_t col = (uint32_t)(__ballot((w >> k) & 1u) >> gshift) & 0xffffu;
        out = k == i ? col : out;
    }
    return out;
}

template <bool FINITE_T, class P>
__global__ __launch_bounds__(kSquareThreads) void square_kernel(CubeJob job)
{
    using Rec = typename P::Rec;
    __shared__ Rec s_delta[17 * kSquareThreads];
    __shared__ uint16_t s_hp[17 * kSquareThreads];
    __shared__ ExpEntry s_exp[kLdsTableEntries];
    const int t = threadIdx.x, g = t >> 4, i = t & 15, gshift = 16 * (g & 3);
    load_exp_table(s_exp, t, job.eps, job.thres);
    __syncthreads(); // the only workgroup barrier: groups run independently from here
    const LineLdsT<Rec, kSquareThreads> S(s_delta, s_hp, t);
    const DecodeCfg cfg{job.tmax, job.thres, job.sched, job.L, s_exp, kLogTable, job.eps, job.eps_log};
    double *le = job.le_scratch + (size_t)blockIdx.x * kCube + (size_t)kSquarePos * g;

    for (size_t sq = (size_t)blockIdx.x * 16 + g; sq < job.n_cw; sq += (size_t)gridDim.x * 16) {
        const double *lch = job.llr + sq * kSquarePos;
        double *soft = job.soft ? job.soft + sq * kSquarePos : nullptr;
        uint32_t half = 0, conv = 0, w = 0, tw = 0;
        uint64_t ng = 0, ngp = 0;
        for (uint32_t it = 1; it <= job.Imax && !conv; ++it) {
            for (int ax = 0; ax < 2; ++ax) {
                ++half; // n_iter += 0.5 before each phase (:145, :184)
                CubeSrc src{lch, le, job.alpha, ax == 0 ? 16 * i : i, ax == 0 ? 1 : 16,
                            it == 1 && ax == 0};
                uint32_t hd, s0;
                double pm_hd;
                int32_t ic;
                LineList ls;
                line_setup_p<P>(src, S, hd, s0, pm_hd, ic, cfg);
                line_search<FINITE_T, P>(S, hd, s0, pm_hd, ic, cfg, ls);
                const double pnotin = fmax(ls.pnl, cfg.eps);
                ListEntries ent;
                list_entries<P>(ls, S, hd, pm_hd, ic, cfg, ent);
                w = 0;
#pragma unroll 1
                for (int j = 0; j < 16; ++j) {
                    const int p = src.base + j * src.stride;
                    const double xin = src(j);
                    double la, lenew;
                    if (ls.nl == 0) { // empty list: L_APP = llr, L_E = 0 (:537-539)
                        la = xin;
                        lenew = 0.0;
                    } else {
                        la = soft_app(ent, ls.nl, pnotin, j, xin, cfg);
                        lenew = la - xin;
                    }
                    // within a phase each position belongs to one line, and a
                    // lane reads a position (src) before it overwrites it
                    le[p] = lenew;
                    if (soft) soft[p] = la;
                    if (!(la > 0.0)) w |= 1u << j;
                }
                // the phase's L_E stores before the next phase's loads (other
                // lanes of the same group read them)
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
                // NG += N_guess; NG_p += the phase's maximum (:162-163, :170)
                uint32_t sum = ls.T, mx = ls.T;
#pragma unroll
                for (int o = 8; o >= 1; o >>= 1) {
                    sum += (uint32_t)__shfl_xor((int)sum, o, 64);
                    const uint32_t m2 = (uint32_t)__shfl_xor((int)mx, o, 64);
                    mx = m2 > mx ? m2 : mx;
                }
                ng += sum;
                ngp += mx;
                // early termination: every line of both axes is a component
                // codeword (the re-encode test of :317-391 for two axes)
                tw = group16_transpose(w, i, gshift);
                const bool ok = syndrome16(w) == 0 && syndrome16(tw) == 0;
                if (((__ballot(!ok) >> gshift) & 0xffffull) == 0) {
                    conv = 1;
                    w = ax == 0 ? w : tw; // rows: lane i holds row a1 = i
                    break;
                }
                if (ax == 1) w = tw;
            }
        }
        // w holds row a1 = i of the final hard decision (bit a0); info bit
        // b = a0 + 8 a1, MSB first (:267-280)
        if (i < kSquareInfoBytes) job.info[sq * kSquareInfoBytes + i] = (uint8_t)rev8(w & 0xffu);
        if (i == 0 && job.stats) {
            CwStatsDev *st = reinterpret_cast<CwStatsDev *>(job.stats) + sq;
            st->half_iters = half;
            st->converged = conv;
            st->NG = ng;
            st->NG_p = ngp;
        }
    }
}

template <bool FINITE_T, class P>
__global__ __launch_bounds__(256) void lines_kernel(const double *llr, size_t n, double *la,
                                                    double *le, uint32_t *nguess, DecodeCfg cfg)
{
    using Rec = typename P::Rec;
    __shared__ Rec s_delta[17 * 256];
    __shared__ uint16_t s_hp[17 * 256];
    __shared__ ExpEntry s_exp[kLdsTableEntries];
    load_exp_table(s_exp, threadIdx.x, cfg.eps, cfg.thres);
    __syncthreads();
    cfg.exptab = s_exp;
    cfg.logtab = kLogTable;
    const size_t line = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (line >= n) return;
    const LineLdsT<Rec, 256> S(s_delta, s_hp, (int)threadIdx.x);
    const double *x = llr + line * 16;
    auto src = [x](int i) { return x[i]; };
    uint32_t hd, s0;
    double pm_hd;
    int32_t ic;
    LineList ls;
    line_setup_p<P>(src, S, hd, s0, pm_hd, ic, cfg);
    line_search<FINITE_T, P>(S, hd, s0, pm_hd, ic, cfg, ls);
    const double pnotin = fmax(ls.pnl, cfg.eps);
    ListEntries ent;
    list_entries<P>(ls, S, hd, pm_hd, ic, cfg, ent);
#pragma unroll 1
    for (int i = 0; i < 16; ++i) {
        const double xi = x[i];
        double a, e;
        if (ls.nl == 0) {
            a = xi;
            e = 0.0;
        } else {
            a = soft_app(ent, ls.nl, pnotin, i, xi, cfg);
            e = a - xi;
        }
        if (la) la[line * 16 + i] = a;
        if (le) le[line * 16 + i] = e;
    }
    if (nguess) nguess[line] = ls.T;
}

// One 256-thread workgroup per codeword: hash the info bits, encode the cube
// in LDS, then each thread writes 16 LLRs (8 Box-Muller pairs).
__global__ __launch_bounds__(256) void gen_kernel(uint64_t seed, uint64_t cw_offset, size_t n_cw,
                                                  const double *sigma, size_t n_sigma, double *llr)
{
    __shared__ uint8_t s_info[kInfoBytes];
    __shared__ uint32_t s_w[256];
    const int t = threadIdx.x;
    for (size_t c = blockIdx.x; c < n_cw; c += gridDim.x) {
        const uint64_t cw = cw_offset + c;
        if (t < 8) {
            const uint64_t w = info_word(seed, cw, t);
#pragma unroll
            for (int q = 0; q < 8; ++q) s_info[8 * t + q] = (uint8_t)(w >> (8 * q));
        }
        __syncthreads();
        if (t < 64) {
            const int i1 = t & 7, i2 = t >> 3;
            s_w[i1 + 16 * i2] = encode_row(rev8(s_info[i1 + 8 * i2]));
        }
        __syncthreads();
        if (t < 64) {
            const int a1 = 8 + (t & 7), i2 = t >> 3;
            uint32_t acc = 0;
            for (int i1 = 0; i1 < 8; ++i1)
                if ((hcol_of((unsigned)i1) >> (a1 - 8)) & 1u) acc ^= s_w[i1 + 16 * i2];
            s_w[a1 + 16 * i2] = acc;
        }
        __syncthreads();
        if (t < 128) {
            const int a1 = t & 15, a2 = 8 + (t >> 4);
            uint32_t acc = 0;
            for (int i2 = 0; i2 < 8; ++i2)
                if ((hcol_of((unsigned)i2) >> (a2 - 8)) & 1u) acc ^= s_w[a1 + 16 * i2];
            s_w[a1 + 16 * a2] = acc;
        }
        __syncthreads();
        const double sg = sigma[n_sigma == 1 ? 0 : c];
        const double var = sg * sg;
        const uint32_t word = s_w[t];
        double *out = llr + c * kCube + 16 * t;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            double z0, z1;
            normal_pair(seed, cw, (uint64_t)(8 * t + q), &z0, &z1);
            const int p0 = 2 * q, p1 = 2 * q + 1;
            const double x0 = ((word >> p0) & 1u) ? -1.0 : 1.0;
            const double x1 = ((word >> p1) & 1u) ? -1.0 : 1.0;
            const double n0 = sg * z0, n1 = sg * z1;
            const double y0 = x0 + n0, y1 = x1 + n1;
            out[p0] = 2.0 * y0 / var;
            out[p1] = 2.0 * y1 / var;
        }
        __syncthreads();
    }
}

__global__ void fill_u32_kernel(uint32_t *p, uint32_t v, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// ---------------------------------------------------------------------------

size_t plain_scratch_slots(size_t n_cw)
{
    const size_t cap = 131072;
    return n_cw < cap ? n_cw : cap;
}

hipError_t launch_plain(const CubeJob &job, int grid, hipStream_t st)
{
    if (job.n_cw == 0) return hipSuccess;
    const bool finite = tmax_binds(job.tmax);
    const dim3 g(grid), b(kPlainThreads);
#define SG_PLAIN(FT, PP) hipLaunchKernelGGL((plain_kernel<FT, PP>), g, b, 0, st, job)
    if (job.arith == kArithF32) {
        if (finite) SG_PLAIN(true, PrecF32); else SG_PLAIN(false, PrecF32);
    } else if (job.arith == kArithF64Legacy) {
        if (finite) SG_PLAIN(true, PrecF64Legacy); else SG_PLAIN(false, PrecF64Legacy);
    } else {
        if (finite) SG_PLAIN(true, PrecF64); else SG_PLAIN(false, PrecF64);
    }
#undef SG_PLAIN
    return hipGetLastError();
}

size_t square_blocks(size_t n_sq)
{
    const size_t cap = 8192; // 131,072 squares in flight
    const size_t b = (n_sq + 15) / 16;
    return b < cap ? b : cap;
}

hipError_t launch_square(const CubeJob &job, int grid, hipStream_t st)
{
    if (job.n_cw == 0) return hipSuccess;
    const bool finite = tmax_binds(job.tmax);
    const dim3 g(grid), b(kSquareThreads);
#define SG_SQUARE(FT, PP) hipLaunchKernelGGL((square_kernel<FT, PP>), g, b, 0, st, job)
    if (job.arith == kArithF32) {
        if (finite) SG_SQUARE(true, PrecF32); else SG_SQUARE(false, PrecF32);
    } else if (job.arith == kArithF64Legacy) {
        if (finite) SG_SQUARE(true, PrecF64Legacy); else SG_SQUARE(false, PrecF64Legacy);
    } else {
        if (finite) SG_SQUARE(true, PrecF64); else SG_SQUARE(false, PrecF64);
    }
#undef SG_SQUARE
    return hipGetLastError();
}

hipError_t launch_lines(const double *llr, size_t n_lines, double *la, double *le,
                        uint32_t *nguess, const uint16_t *sched, uint64_t tmax,
                        double thres, uint32_t L, double eps, double eps_log, int arith,
                        hipStream_t st)
{
    if (n_lines == 0) return hipSuccess;
    const DecodeCfg cfg{tmax, thres, sched, L, nullptr, nullptr, eps, eps_log};
    const dim3 g((unsigned)((n_lines + 255) / 256)), b(256);
    const bool finite = tmax_binds(tmax);
#define SG_LINES(FT, PP)                                                                      \
    hipLaunchKernelGGL((lines_kernel<FT, PP>), g, b, 0, st, llr, n_lines, la, le, nguess, cfg)
    if (arith == kArithF32) {
        if (finite) SG_LINES(true, PrecF32); else SG_LINES(false, PrecF32);
    } else if (arith == kArithF64Legacy) {
        if (finite) SG_LINES(true, PrecF64Legacy); else SG_LINES(false, PrecF64Legacy);
    } else {
        if (finite) SG_LINES(true, PrecF64); else SG_LINES(false, PrecF64);
    }
#undef SG_LINES
    return hipGetLastError();
}

hipError_t launch_gen(uint64_t seed, uint64_t cw_offset, size_t n_cw, const double *sigma,
                      size_t n_sigma, double *llr, hipStream_t st)
{
    if (n_cw == 0) return hipSuccess;
    const unsigned grid = (unsigned)(n_cw < 65536 ? n_cw : 65536);
    hipLaunchKernelGGL(gen_kernel, dim3(grid), dim3(256), 0, st, seed, cw_offset, n_cw, sigma,
                       n_sigma, llr);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t *p, uint32_t v, size_t n, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    unsigned grid = (unsigned)((n + 255) / 256);
    if (grid > 4096) grid = 4096;
    hipLaunchKernelGGL(fill_u32_kernel, dim3(grid), dim3(256), 0, st, p, v, n);
    return hipGetLastError();
}

} // namespace sg
