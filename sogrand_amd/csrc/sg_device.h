// Device-side building blocks of the SOGRAND line decoder (gfx950, wave64).
//
// One lane decodes one (16,8) line. Per-line state lives in registers; the
// per-position path-metric increments and H columns, which the pattern loop
// reads at data-dependent sorted positions, live in LDS in a position-major,
// slot-minor layout (record (b, slot) at index b * nslot + slot), so that lanes
// reading the same position hit consecutive addresses and different positions
// never collide on a bank.
//
// Numerics follow SOGRAND_C/cubic_decoder_final.c exactly (citations are to
// that file): every double is formed by the same operations in the same order;
// the file is compiled with -ffp-contract=off so nothing is fused.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sg_common.h"

#pragma clang fp contract(off)

namespace sg {

struct DecodeCfg {
    uint64_t tmax;         // 0 already mapped to 0x7fffffff by the host
    double thres;          // 1 - p_ET
    const uint16_t *sched; // kSchedTables x kSchedStride masks
    uint32_t L;            // list size 1..kMaxL
};

// Per-lane LDS record arrays: delta (f64) and (H column | natural position << 8).
struct LineLds {
    double *delta;
    uint16_t *hp;
    int nslot;
    int slot;
    __device__ double &d(int b) const { return delta[b * nslot + slot]; }
    __device__ uint16_t &h(int b) const { return hp[b * nslot + slot]; }
};

// JacLog (:762).
__device__ __forceinline__ double jaclog(double x)
{
    if (x > 50.0) return x;
    if (x < -50.0) return 0.0;
    return log(1.0 + exp(x));
}

// x86-64 cvttsd2si into a 32-bit register: NaN / out of range -> INT32_MIN.
__device__ __forceinline__ int32_t cvtt_i32(double r)
{
    if (r >= -2147483648.0 && r < 2147483648.0) return (int32_t)r;
    return INT32_MIN;
}

// Ascending sort of the 16 magnitudes held in the slot's LDS records (d = key,
// h = payload) with the reference's tie order (QuickSort, :763). Used only for
// lines whose magnitudes contain a tie. Segments still to be partitioned are
// kept as bit masks (starts / pending) instead of a recursion stack:
// partitions of disjoint segments commute, so the processing order does not
// change the result.
__device__ __noinline__ void quicksort_lds(const LineLds S)
{
    uint32_t starts = 1u | (1u << 16), pending = 1u;
    while (pending) {
        const int lo = __builtin_ctz(pending);
        pending &= pending - 1u;
        const int hi = __builtin_ctz(starts & ~((2u << lo) - 1u));
        const int len = hi - lo;
        if (len < 2) continue;
        const double pivot = S.d(lo + len / 2);
        int i = 0, j = len - 1;
        while (i <= j) {
            while (S.d(lo + i) < pivot) ++i;
            while (S.d(lo + j) > pivot) --j;
            if (i <= j) {
                const double t = S.d(lo + i);
                S.d(lo + i) = S.d(lo + j);
                S.d(lo + j) = t;
                const uint16_t tp = S.h(lo + i);
                S.h(lo + i) = S.h(lo + j);
                S.h(lo + j) = tp;
                ++i;
                --j;
            }
        }
        if (j >= 0) starts |= 1u << (lo + j + 1);
        if (i < len) starts |= 1u << (lo + i);
        if (j > 0) pending |= 1u << lo;
        if (len - i >= 2) pending |= 1u << (lo + i);
    }
}

// Result of the list search of one line.
struct LineList {
    uint32_t T;         // guesses (:625, :670, :711)
    int nl;             // list entries
    double e[kMaxL];    // exp(-PM_j) of each entry
    uint32_t cw[kMaxL]; // entry codewords, natural order
    double pnl;         // P_neg of the last insertion (getLConf side effect, :772)
};

__device__ __forceinline__ void list_put(LineList &ls, double e, uint32_t cw)
{
    // explicit selects keep the list in registers (no dynamic indexing)
#pragma unroll
    for (int q = 0; q < kMaxL; ++q)
        if (ls.nl == q) {
            ls.e[q] = e;
            ls.cw[q] = cw;
        }
}

// Line set-up (sogrand_main_logic :598-622): hard decision, PM of the hard
// decision in natural order, sort, intercept, and the sorted LDS records.
// Only the sort network holds the 16 magnitudes in registers; the
// transcendental passes are rolled loops over the slot's LDS records, which
// keeps the register footprint of the set-up below that of the search.
// Returns false if an input is not finite.
template <class Src>
__device__ __forceinline__ bool line_setup(const Src &src, const LineLds &S, uint32_t &hd,
                                           uint32_t &s0, double &pm_hd, int32_t &ic)
{
    bool finite = true;
    hd = 0;
    s0 = 0;
    // gather in a rolled loop (magnitudes staged in the slot's LDS records):
    // an unrolled gather would hold all 16 loads and their addresses at once
#pragma unroll 1
    for (int i = 0; i < 16; ++i) {
        const double x = src(i);
        finite = finite && __builtin_isfinite(x);
        if (!(x > 0.0)) { // HardDec (:758)
            hd |= 1u << i;
            s0 ^= hcol_of((unsigned)i);
        }
        S.d(i) = fabs(x);
    }
    if (!finite) return false;
    // getPM_HD: sum of JacLog(-|llr_i|) in natural order, before the sort (:617)
    pm_hd = 0.0;
#pragma unroll 1
    for (int i = 0; i < 16; ++i) pm_hd += jaclog(-S.d(i));
    double a[16];
    uint32_t idx[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        a[i] = S.d(i);
        idx[i] = (uint32_t)hcol_of(i) | ((uint32_t)i << 8);
    }
    // 63-comparator network; exact for distinct keys
#pragma unroll
    for (int c = 0; c < kSortNet.n; ++c) {
        const int lo = kSortNet.lo[c], hi = kSortNet.hi[c];
        const bool sw = a[lo] > a[hi];
        const double al = sw ? a[hi] : a[lo], ah = sw ? a[lo] : a[hi];
        const uint32_t il = sw ? idx[hi] : idx[lo], ih = sw ? idx[lo] : idx[hi];
        a[lo] = al;
        a[hi] = ah;
        idx[lo] = il;
        idx[hi] = ih;
    }
    bool tie = false;
#pragma unroll
    for (int b = 0; b < 15; ++b) tie = tie || (a[b] == a[b + 1]);
    // IC = max(round(absL[0]/beta - 1), 0), beta = (absL[7]-absL[0])/7 (:619-622);
    // the two order statistics do not depend on how ties are arranged
    const double beta = (a[7] - a[0]) / 7.0;
    const int32_t r = cvtt_i32(round(a[0] / beta - 1.0));
    ic = r > 0 ? r : 0;
    if (!tie) {
#pragma unroll
        for (int b = 0; b < 16; ++b) {
            S.d(b) = a[b];
            S.h(b) = (uint16_t)idx[b];
        }
    } else {
        // equal magnitudes: reproduce the reference's quicksort tie order;
        // S.d still holds the natural-order magnitudes
#pragma unroll
        for (int i = 0; i < 16; ++i) S.h(i) = (uint16_t)(hcol_of(i) | (i << 8));
        quicksort_lds(S);
    }
    // delta_b = JacLog(a_b) - JacLog(-a_b): the increment getPM adds (:765)
#pragma unroll 1
    for (int b = 0; b < 16; ++b) {
        const double ab = S.d(b);
        S.d(b) = jaclog(ab) - jaclog(-ab);
    }
    return true;
}

// Record a list entry and apply the stopping rule (:628-641, :674-686, getLConf :767-774).
__device__ __forceinline__ bool list_hit(LineList &ls, double e, uint32_t cw, double pnot,
                                         const DecodeCfg &cfg)
{
    list_put(ls, e, cw);
    double ppos = 0.0;
#pragma unroll
    for (int q = 0; q < kMaxL; ++q)
        if (q <= ls.nl) ppos += ls.e[q];
    double pneg = 0.00390625 * pnot; // 2^-(n-k) * P_notGuess
    ls.pnl = pneg;
    double conf = (ppos + pneg > 1e-40) ? ppos / (ppos + pneg) : 1.0;
    ls.nl += 1;
    return conf > cfg.thres || (uint32_t)ls.nl == cfg.L;
}

// State of one line's guessing loop (sogrand_main_logic :624-744), so that
// the loop can be run in bounded steps.
struct SearchState {
    uint32_t T;       // guesses so far
    uint32_t t;       // next schedule index
    double pnot;      // P_notGuess (:614, :626, :672, :712)
    bool done;
    const uint4 *tp;  // schedule table of the line's intercept
    uint4 nxt;        // prefetched chunk
    uint64_t clo, chi;
    int64_t cur_wt;   // weight level of the last pattern (finite Tmax only)
};

// Hard-decision guess (:624-641) and schedule selection.
__device__ __forceinline__ void search_init(SearchState &ss, LineList &ls, uint32_t hd, uint32_t s0,
                                            double pm_hd, int32_t ic, const DecodeCfg &cfg)
{
    ls.nl = 0;
    ls.pnl = 0.0;
#pragma unroll
    for (int q = 0; q < kMaxL; ++q) {
        ls.e[q] = 0.0;
        ls.cw[q] = 0;
    }
    ss.T = 1;
    const double e_hd = exp(-pm_hd);
    ss.pnot = 1.0 - e_hd;
    ss.done = false;
    if (s0 == 0) ss.done = list_hit(ls, e_hd, hd, ss.pnot, cfg);
    const int tbl = ic < kSchedGenericIc ? ic : kSchedGenericIc;
    ss.tp = reinterpret_cast<const uint4 *>(cfg.sched + (size_t)tbl * kSchedStride);
    ss.nxt = ss.tp[0];
    ss.clo = ss.chi = 0;
    ss.t = 0;
    ss.cur_wt = -1;
}

// Up to max_steps further guesses of the schedule (:652-744), streamed 8 masks
// per 16-byte load with the next load issued one chunk ahead.
template <bool FINITE_T>
__device__ __forceinline__ void search_run(SearchState &ss, LineList &ls, const LineLds &S,
                                           uint32_t hd, uint32_t s0, double pm_hd, int32_t ic,
                                           const DecodeCfg &cfg, uint32_t max_steps)
{
#pragma unroll 1
    for (uint32_t step = 0; step < max_steps && !ss.done; ++step) {
        const uint32_t t = ss.t;
        if ((t & 7u) == 0) {
            ss.clo = (uint64_t)ss.nxt.x | ((uint64_t)ss.nxt.y << 32);
            ss.chi = (uint64_t)ss.nxt.z | ((uint64_t)ss.nxt.w << 32);
            ss.nxt = ss.tp[(t >> 3) + 1]; // the table buffer carries one chunk of padding
        }
        const uint32_t k = t & 7u;
        const uint32_t m = (uint32_t)(((k < 4u) ? ss.clo : ss.chi) >> ((k & 3u) * 16u)) & 0xffffu;
        if (m == 0) { // all 65535 patterns tested (wt > wt_max)
            ss.done = true;
            break;
        }
        if (FINITE_T) {
            int sp = 0;
            for (uint32_t r = m; r; r &= r - 1u) sp += __builtin_ctz(r) + 1;
            const int64_t wt = (int64_t)ic * __builtin_popcount(m) + sp;
            if (wt != ss.cur_wt) { // loop guard once per weight level (:652)
                if (!((uint64_t)ss.T < cfg.tmax)) {
                    ss.done = true;
                    break;
                }
                ss.cur_wt = wt;
            }
        }
        ss.t = t + 1;
        double pm = pm_hd;
        uint32_t syn = s0;
        for (uint32_t r = m; r; r &= r - 1u) {
            const int b = __builtin_ctz(r);
            pm += S.d(b);
            syn ^= S.h(b);
        }
        ss.T += 1;
        const double e = exp(-pm);
        ss.pnot -= e;
        if ((syn & 0xffu) == 0) {
            uint32_t cw = hd;
            for (uint32_t r = m; r; r &= r - 1u) cw ^= 1u << (S.h(__builtin_ctz(r)) >> 8);
            ss.done = list_hit(ls, e, cw, ss.pnot, cfg);
        }
    }
}

template <bool FINITE_T>
__device__ __forceinline__ void line_search(const LineLds &S, uint32_t hd, uint32_t s0,
                                            double pm_hd, int32_t ic, const DecodeCfg &cfg,
                                            LineList &ls)
{
    SearchState ss;
    search_init(ss, ls, hd, s0, pm_hd, ic, cfg);
    search_run<FINITE_T>(ss, ls, S, hd, s0, pm_hd, ic, cfg, 0xffffffffu);
    ls.T = ss.T;
}

// Soft output of position i (SOGRAND_bitSO :536-575) given the line input x.
__device__ __forceinline__ double soft_app(const LineList &ls, double pnotin, int i, double x)
{
    const double eps = 1e-40;
    double pp1 = 1.0 / (1.0 + exp(x));
    double pp0 = 1.0 - pp1;
    pp1 = fmin(fmax(pp1, eps), 1.0 - eps);
    pp0 = fmin(fmax(pp0, eps), 1.0 - eps);
    double p1 = 0.0, p0 = 0.0;
#pragma unroll
    for (int j = 0; j < kMaxL; ++j)
        if (j < ls.nl) {
            double c = (double)((ls.cw[j] >> i) & 1u);
            p1 += c * ls.e[j];
            p0 += (1.0 - c) * ls.e[j];
        }
    p0 += pnotin * pp0;
    p1 += pnotin * pp1;
    return log(fmax(p0, eps)) - log(fmax(p1, eps));
}

} // namespace sg
