// lanegroup_kernel: the latency form of the cube decoder (gfx950). One
// 1024-lane workgroup decodes one cube at a time with the cube resident in LDS:
// the channel LLRs and the extrinsic cube L_E (2 x 32 KB) stay in LDS from the
// cube's first phase to its last, and each phase reads its lines straight out
// of the LDS cube along the phase's axis, so the exchange of extrinsics between
// the three axes is an LDS transpose and the cube costs HBM only its 32 KB of
// LLRs in and 64 B of bits out. Each of the 256 lines of a phase is decoded
// by a group of four lanes (a DPP quad):
//   * set-up: lane g of the quad takes positions g, g + 4, g + 8, g + 12
//     (gather, hard decision, |x|, both JacLogs, rank by counting); the
//     quad's hard decision, tie test and path metric of the hard decision
//     (natural order, :617) are gathered by quad broadcasts;
//   * guessing: the schedule is taken eight patterns at a time, lane g
//     forming the path metrics, syndromes and exp(-pm) of patterns t + 2g and
//     t + 2g + 1; the eight guesses are then accounted in schedule order by every lane of the
//     quad (P_notGuess -= exp(-pm) as the same ordered chain, the list
//     insertion and stopping rule, :652-744), so a line's search takes a
//     quarter of the steps of one lane per line;
//   * soft output: lanes 0..L-1 rebuild one list entry each, broadcast to the
//     quad, and lane g writes positions g + 4k.
// The arithmetic is the line decoder's (sg_device.h) operation for operation,
// so the results are those of the other backends bit for bit. A phase lasts
// as long as its slowest line; with a quarter of the steps per line a cube
// finishes several times sooner than on one lane per line, which is what a
// caller waiting for a small frame sees. Throughput belongs to the persistent
// kernel. Double precision with the final decoder's rules (PrecF64) only.
#include <hip/hip_runtime.h>

#include "sg_common.h"
#include "sg_device.h"
#include "sg_launch.h"

#pragma clang fp contract(off)

namespace sg {

namespace {

constexpr int kLgThreads = 1024;
constexpr int kLgLines = 256;
constexpr int kLgRowStride = kLgLines + 1; // records: one padding record per row

// Lane J of each quad to all four lanes (DPP quad_perm [J, J, J, J]).
template <int J> __device__ __forceinline__ uint32_t qb(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, J * 0x55, 0xf, 0xf, false);
}
template <int J> __device__ __forceinline__ double qb_d(double v)
{
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = qb<J>((uint32_t)u), hi = qb<J>((uint32_t)(u >> 32));
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
// OR and max over the quad (swap neighbours, then swap pairs).
__device__ __forceinline__ uint32_t quad_or(uint32_t v)
{
    v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xb1, 0xf, 0xf, false);
    v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4e, 0xf, 0xf, false);
    return v;
}
__device__ __forceinline__ uint32_t quad_max(uint32_t v)
{
    uint32_t w = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xb1, 0xf, 0xf, false);
    v = w > v ? w : v;
    w = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4e, 0xf, 0xf, false);
    return w > v ? w : v;
}
// Line (lo, hi) = (line & 15, line >> 4) of axis ax covers base + stride i:
// axis 0 runs over a0 (lines (a1, a2)), axis 1 over a1 (lines (a0, a2)),
// axis 2 over a2 (lines (a0, a1)), idx = a0 + 16 a1 + 256 a2 (final.c:157-159,
// :196-198, :235-237).
struct LineAxis {
    int base, stride;
};
__device__ __forceinline__ LineAxis line_axis(int ax, int line)
{
    const int lo = line & 15, hi = line >> 4;
    if (ax == 0) return {16 * lo + 256 * hi, 1};
    if (ax == 1) return {lo + 256 * hi, 16};
    return {lo + 16 * hi, 256};
}

// The four bits of a wave ballot that belong to this lane's quad.
__device__ __forceinline__ uint32_t quad_bits(bool pred, int lane)
{
    return (uint32_t)(__ballot(pred) >> (lane & 60)) & 15u;
}

// Bits past the sixth of a pattern (position word w has kPosMore; about one
// guess in 16,000 at 2 dB): their increments and H columns from the mask.
template <class LDS>
__device__ __forceinline__ void more_bits(const LDS &S, uint32_t w, const uint16_t *masks,
                                          uint32_t t, double &pm, uint32_t &syn)
{
    if (!(w & kPosMore)) return;
    uint32_t m = masks[t];
#pragma unroll 1
    for (uint32_t i = 0; i < kPosFields; ++i) m &= m - 1u;
#pragma unroll 1
    for (; m; m &= m - 1u) {
        const int b = __builtin_ctz(m);
        syn ^= (uint32_t)S.h(b);
        pm = pm + S.d(b);
    }
}

} // namespace

template <bool FINITE_T>
__global__ __launch_bounds__(kLgThreads) void lanegroup_kernel(CubeJob job)
{
    using P = PrecF64;
    __shared__ double s_lch[kCube];
    __shared__ double s_le[kCube];
    __shared__ double s_d[17 * kLgRowStride];
    __shared__ uint16_t s_h[17 * kLgRowStride];
    __shared__ uint8_t s_hd[kCube];
    __shared__ double s_ord[kLgLines][2]; // |x| of ranks 0 and 7 per line
    __shared__ uint32_t s_sum[kLgThreads / 64], s_max[kLgThreads / 64], s_bad[kLgThreads / 64];
    __shared__ ExpEntry s_exp[kLdsTableEntries];

    const int t = threadIdx.x, lane = t & 63;
    load_exp_table(s_exp, t, job.eps, job.thres);
    const DecodeCfg cfg{job.tmax, job.thres, job.sched, job.L, s_exp, kLogTable, job.eps, job.eps_log};
    if ((t & 3) == 0) {
        // null records read by empty position-word fields (row 0: nothing
        // else writes it)
        const LineLdsT<double, kLgLines, kLgRowStride> S0(s_d, s_h, t >> 2);
        S0.d(-1) = 0.0;
        S0.h(-1) = 0;
    }

    for (size_t cube = blockIdx.x; cube < job.n_cw; cube += gridDim.x) {
        const double *lch = job.llr + cube * kCube;
        double *soft = job.soft ? job.soft + cube * kCube : nullptr;
        // (the thread index and the zero made opaque: formed here, not held
        // in registers for the whole kernel)
        int tl = t;
        double zero = 0.0;
        asm volatile("" : "+v"(tl), "+v"(zero));
#pragma unroll
        for (int k = 0; k < kCube / kLgThreads; ++k) {
            const int p = tl + k * kLgThreads;
            s_lch[p] = lch[p];
            s_le[p] = zero; // L_E = 0 before the first phase (:129)
        }
        __syncthreads();
        uint32_t half = 0, conv = 0;
        uint64_t ng = 0, ngp = 0;
        for (uint32_t it = 1; it <= job.Imax && !conv; ++it) {
            for (int ax = 0; ax < 3; ++ax) {
                ++half; // n_iter += 0.5 before each phase (:145, :184, :223)
                // the line index is made opaque per phase, so that the
                // compiler forms the line's LDS addresses here instead of
                // keeping all of them live (in spilled registers) for the cube
                int line = t >> 2, g = t & 3;
                asm volatile("" : "+v"(line), "+v"(g));
                const LineLdsT<double, kLgLines, kLgRowStride> S(s_d, s_h, line);
                const LineAxis pg = line_axis(ax, line);
                // ---- set-up (sogrand_main_logic :598-622), positions g + 4k ----
                double a[4];
                uint32_t negbits = 0;
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int i = g + 4 * k;
                    const int p = pg.base + i * pg.stride;
                    const double tt = job.alpha * s_le[p]; // L_A = alpha L_E (:147-155)
                    const double x = s_lch[p] + tt;
                    if (!(x > 0.0)) negbits |= 1u << i; // HardDec (:758)
                    a[k] = fabs(x);
                    S.d(i) = a[k];
                }
                const uint32_t hd = quad_or(negbits);
                const uint32_t s0 = syndrome16(hd);
                // rank by counting: the number of smaller magnitudes, a
                // permutation exactly when the magnitudes are distinct
                uint32_t rank[4] = {0, 0, 0, 0};
#pragma unroll 4
                for (int j = 0; j < 16; ++j) {
                    const double aj = S.d(j);
#pragma unroll
                    for (int k = 0; k < 4; ++k) rank[k] += aj < a[k] ? 1u : 0u;
                }
                uint32_t seen = 0;
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    seen |= 1u << rank[k];
                    if (rank[k] == 0) s_ord[line][0] = a[k];
                    if (rank[k] == 7) s_ord[line][1] = a[k];
                }
                const bool tie = quad_or(seen) != 0xffffu;
                double pm_hd = 0.0;
                int32_t ic = 0;
                if (!tie) {
                    ic = intercept(s_ord[line][0], s_ord[line][1], false); // :619-622
                    // getPM_HD (:617) in natural order (position i is slot i / 4
                    // of lane i % 4) and getPM's increments (:765), each written
                    // straight to its sorted record: the keys are all ranked
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        const int i = g + 4 * k;
                        const double jn = jaclog(-a[k], P::kJacLim, cfg);
                        const double delta = jaclog(a[k], P::kJacLim, cfg) - jn;
                        pm_hd += qb_d<0>(jn);
                        pm_hd += qb_d<1>(jn);
                        pm_hd += qb_d<2>(jn);
                        pm_hd += qb_d<3>(jn);
                        S.d((int)rank[k]) = delta;
                        S.h((int)rank[k]) = (uint16_t)(hcol_of((unsigned)i) | (i << 8));
                    }
                } else {
                    // equal magnitudes or a NaN: the reference's quicksort
                    // decides their order; one lane replays it on the records
                    // (still |x| in natural order) and the quad takes its PM_HD
                    // and intercept
                    double pmh = 0.0;
                    int32_t icq = 0;
                    if (g == 0) {
                        double a_first, a_eighth;
                        pmh = setup_with_ties(S, P::kJacLim, cfg, a_first, a_eighth);
                        icq = intercept(a_first, a_eighth, false);
                    }
                    pm_hd = qb_d<0>(pmh);
                    ic = (int32_t)qb<0>((uint32_t)icq);
                }
                // ---- guessing (:624-744), eight patterns per step: lane g
                // takes schedule positions t + 2g and t + 2g + 1 (two
                // independent metric and exp chains per lane) ----
                SearchState ss;
                LineList ls;
                search_start<P>(ss, ls, s0, pm_hd, cfg);
                const int tbl = ic < kSchedGenericIc ? ic : kSchedGenericIc;
                const uint16_t *masks = cfg.sched + (size_t)tbl * kSchedStride;
                const uint2 *wpairs = reinterpret_cast<const uint2 *>(
                    pos_tables(cfg.sched) + (size_t)tbl * kSchedStride);
                uint32_t pm_hi = 0;
                uint2 wnext = wpairs[g];
#pragma unroll 1
                while (__ballot(!ss.done) != 0ull) {
                    if (!ss.done) {
                        const uint32_t t0 = ss.t; // a multiple of 8
                        const uint2 w = wnext;
                        // the next step's words (the tables carry padding past
                        // their end marker, the last one up to index 65543)
                        wnext = wpairs[(t0 >> 1) + 4u + (uint32_t)g];
                        const uint32_t ta = t0 + 2u * (uint32_t)g;
                        double pm1, pm2;
                        uint32_t syn1, syn2;
                        pattern_sums<P, true>(S, s0, pm_hd, w.x, w.y, pm1, pm2, syn1, syn2);
                        if ((w.x | w.y) & kPosMore) {
                            more_bits(S, w.x, masks, ta, pm1, syn1);
                            more_bits(S, w.y, masks, ta + 1u, pm2, syn2);
                        }
                        const double e1 = exp_neg_loop(pm1, cfg.exptab);
                        const double e2 = exp_neg_loop(pm2, cfg.exptab);
                        const uint32_t h1 = (uint32_t)(__builtin_bit_cast(uint64_t, pm1) >> 32);
                        const uint32_t h2 = (uint32_t)(__builtin_bit_cast(uint64_t, pm2) >> 32);
                        pm_hi = max(pm_hi, max(h1, h2));
                        // a pattern stops the search before it: the end marker,
                        // or a weight level over budget (T < Tmax taken before
                        // each level, :652, with T = t + 1)
                        bool before1 = w.x == 0u, before2 = w.y == 0u;
                        if (FINITE_T) {
                            before1 = before1 || ((w.x & kPosLevel) && !((uint64_t)ta + 1u < cfg.tmax));
                            before2 = before2 || ((w.y & kPosLevel) && !((uint64_t)ta + 2u < cfg.tmax));
                        }
                        const uint32_t bm[2] = {quad_bits(before1, lane), quad_bits(before2, lane)};
                        const uint32_t hm[2] = {quad_bits((syn1 & 0xffu) == 0, lane),
                                                quad_bits((syn2 & 0xffu) == 0, lane)};
                        const uint32_t big[2] = {quad_bits(!(pm1 < 512.0), lane),
                                                 quad_bits(!(pm2 < 512.0), lane)};
                        // the weights in schedule order: pattern j is lane j / 2's
                        // first (j even) or second
                        const double ev[8] = {qb_d<0>(e1), qb_d<0>(e2), qb_d<1>(e1), qb_d<1>(e2),
                                              qb_d<2>(e1), qb_d<2>(e2), qb_d<3>(e1), qb_d<3>(e2)};
                        uint32_t used = 8;
                        bool stop = false;
#pragma unroll
                        for (int j = 0; j < 8; ++j) {
                            if (stop) continue;
                            if ((bm[j & 1] >> (j >> 1)) & 1u) {
                                stop = true;
                                used = (uint32_t)j;
                                continue;
                            }
                            ss.pnot -= ev[j];
                            if ((hm[j & 1] >> (j >> 1)) & 1u) {
                                // a hit whose weight the loop's exp stood in for
                                // poisons P_notGuess (exact re-run below)
                                if ((big[j & 1] >> (j >> 1)) & 1u) ss.pnot = __builtin_nan("");
                                if (list_hit(ls, ev[j], t0 + (uint32_t)j + 1u, ss.pnot, cfg)) {
                                    stop = true;
                                    used = (uint32_t)j + 1u;
                                }
                            }
                        }
                        ss.t = t0 + used;
                        ss.done = stop;
                    }
                }
                // a search whose chain may have needed an exact weight (see
                // exp_neg_loop; every lane of the quad decides alike)
                pm_hi = quad_max(pm_hi);
                if ((!chain_exact(ss.pnot) && pm_hi >= 0x40800000u) || pm_hi > 0x7ff00000u)
                    search_exact<FINITE_T, P>(ss, ls, S, hd, s0, pm_hd, ic, cfg);
                ls.T = ss.t + 1u;
                // ---- soft output (SOGRAND_bitSO :536-575) ----
                double my_e = 0.0;
                uint32_t my_cw = 0;
                if (g < ls.nl) list_entry<P>(ls, g, S, hd, pm_hd, ic, cfg, my_e, my_cw);
                ListEntries ent;
                ent.e[0] = qb_d<0>(my_e);
                ent.e[1] = qb_d<1>(my_e);
                ent.e[2] = qb_d<2>(my_e);
                ent.cw[0] = qb<0>(my_cw);
                ent.cw[1] = qb<1>(my_cw);
                ent.cw[2] = qb<2>(my_cw);
                const double pnotin = fmax(ls.pnl, cfg.eps);
#pragma unroll 1
                for (int k = 0; k < 4; ++k) {
                    const int i = g + 4 * k;
                    const int p = pg.base + i * pg.stride;
                    const double tt = job.alpha * s_le[p];
                    const double xin = s_lch[p] + tt; // the set-up's input
                    double la, lenew;
                    if (ls.nl == 0) { // empty list: L_APP = llr, L_E = 0 (:537-539)
                        la = xin;
                        lenew = 0.0;
                    } else {
                        la = soft_app(ent, ls.nl, pnotin, i, xin, cfg);
                        lenew = la - xin;
                    }
                    s_le[p] = lenew;
                    if (soft) soft[p] = la;
                    s_hd[p] = (la > 0.0) ? 0 : 1;
                }
                // NG += N_guess; NG_p += max over the phase's lines (:162-163,
                // :170): lane 0 of each quad carries its line's count
                const uint32_t tg = g == 0 ? ls.T : 0u;
                uint32_t ws = tg, wm = tg;
#pragma unroll
                for (int o = 32; o >= 1; o >>= 1) {
                    ws += __shfl_xor(ws, o, 64);
                    const uint32_t v = __shfl_xor(wm, o, 64);
                    wm = v > wm ? v : wm;
                }
                if (lane == 0) {
                    int wv = t >> 6;
                    asm volatile("" : "+v"(wv));
                    s_sum[wv] = ws;
                    s_max[wv] = wm;
                }
                __syncthreads();
                uint32_t bs = 0, bmx = 0;
#pragma unroll
                for (int w = 0; w < kLgThreads / 64; ++w) {
                    bs += s_sum[w];
                    bmx = s_max[w] > bmx ? s_max[w] : bmx;
                }
                // the same in every lane: the cube's counts stay in scalar registers
                ng += __builtin_amdgcn_readfirstlane(bs);
                ngp += __builtin_amdgcn_readfirstlane(bmx);
                // early termination: all 768 lines of the hard decision are
                // component codewords (equivalent to the re-encode test :317-391)
                bool line_bad = false;
                if (t < 3 * kLgLines) {
                    int et = t;
                    asm volatile("" : "+v"(et));
                    const LineAxis eg = line_axis(et >> 8, et & 255);
                    uint32_t w = 0;
#pragma unroll
                    for (int j = 0; j < 16; ++j) w |= (uint32_t)s_hd[eg.base + j * eg.stride] << j;
                    line_bad = syndrome16(w) != 0;
                }
                // a workgroup OR by wave ballots (a barrier separates these
                // writes from the reads of the previous phase)
                const bool wave_bad = __ballot(line_bad) != 0ull;
                if (lane == 0) {
                    int wv = t >> 6;
                    asm volatile("" : "+v"(wv));
                    s_bad[wv] = wave_bad ? 1u : 0u;
                }
                __syncthreads();
                uint32_t any_bad = 0;
#pragma unroll
                for (int w = 0; w < kLgThreads / 64; ++w) any_bad |= s_bad[w];
                if (!__builtin_amdgcn_readfirstlane(any_bad)) {
                    conv = 1;
                    break;
                }
            }
        }
        // (the thread index made opaque again: these addresses are formed here)
        int tc = t;
        asm volatile("" : "+v"(tc));
        if (tc < kInfoBytes) { // info bit b = a0 + 8*a1 + 64*a2, MSB first (:267-280)
            uint32_t byte = 0;
            const int a1 = tc & 7, a2 = tc >> 3;
#pragma unroll
            for (int q = 0; q < 8; ++q) byte |= (uint32_t)s_hd[q + 16 * a1 + 256 * a2] << (7 - q);
            job.info[cube * kInfoBytes + tc] = (uint8_t)byte;
        }
        if (t == 0 && job.stats) {
            uint32_t *st = job.stats + cube * 6;
            st[0] = half;
            st[1] = conv;
            st[2] = (uint32_t)ng;
            st[3] = (uint32_t)(ng >> 32);
            st[4] = (uint32_t)ngp;
            st[5] = (uint32_t)(ngp >> 32);
        }
        __syncthreads(); // s_hd and the cube are read above before the next cube
    }
}

size_t lanegroup_blocks(size_t n_cw, int cu_count)
{
    // one resident workgroup per CU (the LDS holds one cube); more workgroups
    // than that only queue
    const size_t cap = (size_t)(cu_count > 0 ? cu_count : 256);
    return n_cw < cap ? (n_cw > 0 ? n_cw : 1) : cap;
}

hipError_t launch_lanegroup(const CubeJob &job, int grid, hipStream_t st)
{
    const dim3 g((unsigned)grid), b(kLgThreads);
    if (tmax_binds(job.tmax))
        hipLaunchKernelGGL((lanegroup_kernel<true>), g, b, 0, st, job);
    else
        hipLaunchKernelGGL((lanegroup_kernel<false>), g, b, 0, st, job);
    return hipGetLastError();
}

} // namespace sg
