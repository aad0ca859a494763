// Device-side building blocks of the SOGRAND line decoder (gfx950, wave64).
//
// One lane decodes one (16,8) line. Per-line state lives in registers; the
// per-position path-metric increments and H columns, which the pattern loop
// reads at data-dependent sorted positions, live in LDS in a position-major,
// slot-minor layout (record (b, slot) at index (b + 1) * NS + slot, row 0 a null
// record), so that lanes reading the same position hit consecutive addresses.
//
// Numerics follow SOGRAND_C/cubic_decoder_final.c exactly (citations are to
// that file unless another is named): every double is formed by the same
// operations in the same order; the file is compiled with -ffp-contract=off so
// nothing is fused. Where the reference calls the C library's exp and log,
// the decoder uses table-driven ones with the library's own method, constants
// and fused operations (sg_exp.h, sg_log.h), which give the library's results
// bit for bit. The policy PrecF64Legacy
// follows the reference's earlier SOGRAND_C/cubic_decoder.c instead.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sg_common.h"
#include "sg_conf.h"
#include "sg_exp.h"
#include "sg_log.h"

#pragma clang fp contract(off)

namespace sg {

struct DecodeCfg {
    uint64_t tmax;          // 0 already mapped to 0x7fffffff by the host
    double thres;           // 1 - p_ET
    const uint16_t *sched;  // schedule buffer: kSchedTables x kSchedStride masks, then
                            // the position words (pos_tables, sg_common.h)
    uint32_t L;             // list size 1..kMaxL
    const ExpEntry *exptab; // exp table in LDS, then {eps, thres} (load_exp_table)
    const LogEntry *logtab; // log table (kLogTable, in global memory: read only by
                            // the set-up and the soft output, it stays in the caches)
    double eps;             // EPSILON of the confidence test, P_notinlist floor and clamps
    double eps_log;         // floor inside the soft output's logs (eps, or 1e-30 for
                            // cubic_decoder.c:533)
};

// Entries of the LDS table: the exp table, then the entries {eps, thres} and
// {conf_gap(thres, eps), 0} that the list insertion reads where it needs them
// (as broadcast loads, which keep the doubles out of the scalar registers the
// search loop runs short of), then the log's polynomial coefficients
// (kLogCoef, two per entry).
constexpr int kLdsGapEntry = kExpTableEntries + 1;
constexpr int kLdsLogCoefEntry = kExpTableEntries + 2;
constexpr int kLdsTableEntries = kLdsLogCoefEntry + kLogCoefs / 2;

// Fill the LDS table (first kLdsTableEntries threads) - the caller synchronises.
__device__ __forceinline__ void load_exp_table(ExpEntry *lds, int tid, double eps, double thres)
{
    if (tid < kExpTableEntries) {
        lds[tid] = exp_table_entry(tid);
    } else if (tid == kExpTableEntries) {
        lds[tid] = ExpEntry{eps, thres};
    } else if (tid == kLdsGapEntry) {
        lds[tid] = ExpEntry{conf_gap(thres, eps), 0.0};
    } else if (tid < kLdsTableEntries) {
        const int j = 2 * (tid - kLdsLogCoefEntry);
        lds[tid] = ExpEntry{kLogCoef[j], kLogCoef[j + 1]};
    }
}

// The log's coefficients in the LDS table, read where the log runs: the table
// is never written after it is filled, so without the opaque zero offset the
// compiler would hoist all fifteen loads into thirty kernel-wide registers.
__device__ __forceinline__ const double *log_coefs(const DecodeCfg &cfg)
{
    int z = 0;
    asm volatile("" : "+v"(z));
    return reinterpret_cast<const double *>(cfg.exptab + kLdsLogCoefEntry + z);
}

// exp(-pm) for pm >= 0, the probability weight of one guess (:626, :672, :712).
// Every lane forms the main path; the rare metrics from 512 on (and +inf) take
// the library's large-argument branch, which runs only if some lane needs it.
// (For -pm above -2^-54 the main path gives the library's 1 + x as well.)
__device__ __forceinline__ double exp_neg(double pm, const ExpEntry *tab)
{
    double y = exp_main(-pm, tab);
    if (!(pm < 512.0)) y = exp_large(-pm, tab);
    return y;
}

// exp(-pm) of a guess in the search loop: the main path, with pm capped at
// 708 (one v_min_f64) so that the scale stays a normal double. Exact for
// pm < 512. A larger metric gets a stand-in in (0, 2^-738] instead of its
// weight (also below 2^-738): subtracting either from a P_notGuess of magnitude
// 2^-680 or more leaves it unchanged (both are below half its ulp), and since
// P_notGuess never increases, a search that ends with P_notGuess >= 2^-680
// has formed the exact chain. A hit with such a metric (whose weight would
// enter the found mass) poisons P_notGuess with NaN, which ends the search at
// once; a search that ends with P_notGuess below 2^-680 or NaN is then run
// again with exact weights (chain_exact, search_exact).
__device__ __forceinline__ double exp_neg_loop(double pm, const ExpEntry *tab)
{
    // the cap from a scalar register (a vector one would be rebuilt per use)
    double pc;
    asm("v_min_f64 %0, %1, %2" : "=v"(pc) : "v"(pm), "s"(708.0));
    return exp_main(-pc, tab);
}

// Whether a finished search formed its chain exactly (see exp_neg_loop).
__device__ __forceinline__ bool chain_exact(double pnot)
{
    return pnot >= 0x1p-680;
}

// exp(x) for any x (the soft output's exp(llr), :547).
__device__ __forceinline__ double exp_any(double x, const ExpEntry *tab)
{
    return exp_of(x, tab);
}

// Arithmetic of the per-guess path metrics.
//   PrecF64: everything in double, bit-exact with the reference.
//   PrecF32: reliabilities sorted as floats, path-metric increments stored and
//            summed as floats, exp(-pm) of each guess in single precision; the
//            1 - found-mass chain (P_notGuess), PM of the hard decision, list
//            weights (exp(-pm) of hits, recomputed in double), confidence and
//            soft output stay in double.
//   PrecF64Legacy: double, with the rules of the reference's earlier decoder
//            SOGRAND_C/cubic_decoder.c: JacLog cut at +-30 (:699), IC = 0 when
//            beta <= 1e-9 (:580-585), and P_notGuess never lowered for this
//            code (:601, :650: only even codes subtract), so the confidence
//            P_pos / (P_pos + 2^-8) stays below 1 - 1e-5 and every search runs
//            until its list is full.
struct PrecF64 {
    using Rec = double;
    static constexpr bool kF32 = false;
    static constexpr bool kLegacy = false;
    static constexpr double kJacLim = 50.0;
};
struct PrecF32 {
    using Rec = float;
    static constexpr bool kF32 = true;
    static constexpr bool kLegacy = false;
    static constexpr double kJacLim = 50.0;
};
struct PrecF64Legacy {
    using Rec = double;
    static constexpr bool kF32 = false;
    static constexpr bool kLegacy = true;
    static constexpr double kJacLim = 30.0;
};

// exp(-pm) of one guess in single precision (v_exp_f32 on -pm log2 e), widened.
__device__ __forceinline__ double exp_neg_f32(float pm)
{
    return (double)__builtin_amdgcn_exp2f(pm * -1.44269504088896341f);
}

// Per-lane LDS record columns of an NS-slot array: delta (Rec) and (H column |
// natural position << 8) of sorted positions b = 0..15 at rows 1..16, and at
// row 0 a null record (delta +0.0, H column 0) that the pattern rounds read
// for the empty fields (row 0) of a position word. The column pointers fold the
// slot in, so a record address is one shift-add of the row. Two 2-byte H
// records share a 4-byte LDS word; the pair that shares one is slots s and
// s + 32 of each group of 64 (column h_column(s)), which sit in different
// halves of a wave, so the lanes of one half never meet on a bank when they
// read different rows. RS is the row stride in records: NS where each lane
// reads its own slot, NS + 1 where several lanes read different rows of one
// slot (one record of padding per row moves each row to the next bank).
__host__ __device__ constexpr int h_column(int slot)
{
    return (slot & ~63) | ((slot & 31) << 1) | ((slot >> 5) & 1);
}
template <class RecT, int NS, int RS = NS> struct LineLdsT {
    static_assert(NS % 64 == 0, "record arrays hold whole waves");
    static_assert(RS >= NS, "rows hold every slot");
    using Rec = RecT;
    Rec *dcol;       // this slot's record of sorted position 0
    uint16_t *hcol;
    __device__ LineLdsT(Rec *delta, uint16_t *hp, int slot)
        : dcol(delta + RS + slot), hcol(hp + RS + h_column(slot))
    {
    }
    __device__ Rec &d(int b) const { return dcol[b * RS]; }
    __device__ uint16_t &h(int b) const { return hcol[b * RS]; }
    // the same records by row (row = b + 1, row 0 the null record), as the
    // fields of a position word give them
    __device__ Rec dr(uint32_t row) const { return dcol[((int)row - 1) * RS]; }
    __device__ uint16_t hr(uint32_t row) const { return hcol[((int)row - 1) * RS]; }
};

// JacLog (:762; cut at lim = 50, or 30 in cubic_decoder.c:699).
__device__ __forceinline__ double jaclog(double x, double lim, const DecodeCfg &cfg)
{
    if (x > lim) return x;
    if (x < -lim) return 0.0;
    return log_pos(1.0 + exp_main(x, cfg.exptab), cfg.logtab, log_coefs(cfg)); // |x| <= lim < 512
}


// x86-64 cvttsd2si into a 32-bit register: NaN / out of range -> INT32_MIN.
__device__ __forceinline__ int32_t cvtt_i32(double r)
{
    if (r >= -2147483648.0 && r < 2147483648.0) return (int32_t)r;
    return INT32_MIN;
}

// IC = max(round(absL[0]/beta - 1), 0), beta = (absL[7]-absL[0])/7 (:619-622);
// cubic_decoder.c:580-585 takes IC = 0 unless beta > 1e-9.
__device__ __forceinline__ int32_t intercept(double a_first, double a_eighth, bool guard)
{
    const double beta = (a_eighth - a_first) / 7.0;
    const int32_t rr = cvtt_i32(round(a_first / beta - 1.0));
    const int32_t ic = rr > 0 ? rr : 0;
    return (guard && !(beta > 1e-9)) ? 0 : ic;
}

// Ascending sort of the 16 magnitudes held in the slot's LDS records (d = key,
// h = payload) with the reference's tie order (QuickSort, :763). Used only for
// lines whose magnitudes contain a tie. Segments still to be partitioned are
// kept as bit masks (starts / pending) instead of a recursion stack:
// partitions of disjoint segments commute, so the processing order does not
// change the result.
template <class LDS> __device__ __noinline__ void quicksort_lds(const LDS S)
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

// Result of the list search of one line. The search keeps only what its
// stopping rule needs: the running found mass of the list (getLConf's P_pos,
// :769, summed in list order, so a running sum gives the same double), the
// last P_neg and, per entry, the 16-bit position of its pattern in the
// schedule (0 for the hard decision, else schedule index + 1). The entries'
// weights and codewords, needed only by the soft output, are rebuilt from
// those positions by list_entry with the same arithmetic.
struct LineList {
    uint32_t T;     // guesses (:625, :670, :711)
    int nl;         // list entries
    double ppos;    // sum of exp(-PM_j) over the entries, in list order
    double pnl;     // P_neg of the last insertion (getLConf side effect, :772)
    uint64_t code;  // 16 bits per entry: schedule position + 1, 0 = hard decision
};

// One list entry: exp(-PM) and the codeword (natural order).
struct ListEntries {
    double e[kMaxL];
    uint32_t cw[kMaxL];
};

// Set-up of a line whose magnitudes contain a tie, from d(i) = |x_i| in natural
// order: the reference's quicksort decides the order among equal magnitudes
// (replayed by quicksort_lds), then PM_HD in natural order and the sorted
// records. Returns PM_HD; a_first and a_eighth receive the smallest and the
// eighth smallest magnitude (the order statistics of IC, :619-621).
template <class LDS>
__device__ __forceinline__ double setup_with_ties(const LDS S, double jlim, const DecodeCfg &cfg,
                                                  double &a_first, double &a_eighth)
{
#pragma unroll 1
    for (int i = 0; i < 16; ++i) S.h(i) = (uint16_t)(hcol_of((unsigned)i) | (i << 8));
    quicksort_lds(S);
    a_first = S.d(0);
    a_eighth = S.d(7);
    // natural-order PM_HD (:617): position i now sits at its sorted slot
    uint64_t slot_of = 0;
#pragma unroll 1
    for (int b = 0; b < 16; ++b) slot_of |= (uint64_t)b << (4 * (S.h(b) >> 8));
    double pm_hd = 0.0;
#pragma unroll 1
    for (int i = 0; i < 16; ++i)
        pm_hd += jaclog(-S.d((int)((slot_of >> (4 * i)) & 15u)), jlim, cfg);
#pragma unroll 1
    for (int b = 0; b < 16; ++b) {
        const double ab = S.d(b);
        S.d(b) = jaclog(ab, jlim, cfg) - jaclog(-ab, jlim, cfg);
    }
    return pm_hd;
}

// Move the natural-order records d(i) to their sorted positions rank_i (4 bits
// per position in `rank`), one permutation cycle at a time, and write the H
// column | natural position records alongside.
template <class LDS> __device__ __forceinline__ void place_by_rank(const LDS &S, uint64_t rank)
{
    using Rec = typename LDS::Rec;
    uint32_t placed = 0;
#pragma unroll 1
    for (int start = 0; start < 16; ++start) {
        if ((placed >> start) & 1u) continue;
        Rec carry = S.d(start);
        int from = start;
#pragma unroll 1
        for (;;) {
            const int to = (int)((rank >> (4 * from)) & 15u);
            placed |= 1u << to;
            S.h(to) = (uint16_t)(hcol_of((unsigned)from) | (from << 8));
            if (to == start) {
                S.d(to) = carry;
                break;
            }
            const Rec t = S.d(to);
            S.d(to) = carry;
            carry = t;
            from = to;
        }
    }
}

// Line set-up (sogrand_main_logic :598-622): hard decision, PM of the hard
// decision in natural order, sort, intercept, and the sorted LDS records.
//
// Everything runs as rolled loops over the slot's 16 LDS records, so the set-up
// holds only a handful of registers (it runs while other lanes of the wave keep
// their search state live):
//   1. gather |x_i| into d(i);
//   2. rank by counting: rank_i = #{j : a_j < a_i} + #{j < i : a_j == a_i};
//      any equal pair is a tie;
//   3. without ties (the usual case) the ranks are the unique ascending order:
//      delta_i = JacLog(a_i) - JacLog(-a_i) is formed in natural order (sharing
//      JacLog(-a_i) with PM_HD) and moved to d(rank_i) by following the cycles
//      of the permutation in place;
//   4. with ties, the reference's quicksort decides the order among equal
//      magnitudes, so it is replayed on the records (quicksort_lds).
// Every double is taken as the reference takes it: an infinite magnitude
// gives an infinite increment, and a NaN flows through the sort (as a tie), the
// intercept and the metrics. legacy: cubic_decoder.c's JacLog cut and IC guard.
template <class Src, class LDS>
__device__ __forceinline__ void line_setup(const Src &src, const LDS &S, uint32_t &hd,
                                           uint32_t &s0, double &pm_hd, int32_t &ic,
                                           bool legacy, const DecodeCfg &cfg)
{
    const double jlim = legacy ? 30.0 : 50.0;
    hd = 0;
    s0 = 0;
#pragma unroll 1
    for (int i = 0; i < 16; ++i) {
        const double x = src(i);
        if (!(x > 0.0)) { // HardDec (:758)
            hd |= 1u << i;
            s0 ^= hcol_of((unsigned)i);
        }
        S.d(i) = fabs(x);
    }

    // rank = the number of smaller magnitudes: a permutation exactly when the
    // magnitudes are distinct and not NaN, so a repeated rank (a missing bit in
    // seen) reveals a tie or a NaN, which the quicksort replay below orders
    uint64_t rank = 0; // 4 bits per natural position
    uint32_t seen = 0;
    double a_first = 0.0, a_eighth = 0.0;
#pragma unroll 1
    for (int i = 0; i < 16; ++i) {
        const double ai = S.d(i);
        uint32_t r = 0;
#pragma unroll 4
        for (int j = 0; j < 16; ++j) r += S.d(j) < ai ? 1u : 0u;
        rank |= (uint64_t)r << (4 * i);
        seen |= 1u << r;
        if (r == 0) a_first = ai;
        if (r == 7) a_eighth = ai;
    }
    const bool tie = seen != 0xffffu;

    pm_hd = 0.0;
    if (!tie) {
        ic = intercept(a_first, a_eighth, legacy);
        // getPM_HD (:617) in natural order; delta_i (getPM's increment, :765)
#pragma unroll 1
        for (int i = 0; i < 16; ++i) {
            const double ai = S.d(i);
            const double jn = jaclog(-ai, jlim, cfg);
            pm_hd += jn;
            S.d(i) = jaclog(ai, jlim, cfg) - jn;
        }
        place_by_rank(S, rank);
    } else {
        pm_hd = setup_with_ties(S, jlim, cfg, a_first, a_eighth);
        ic = intercept(a_first, a_eighth, legacy);
    }
    S.d(-1) = 0.0; // null record read by empty position-word fields
    S.h(-1) = 0;
}

// Set-up of the single-precision variant: reliabilities ranked as floats (equal
// floats ordered by position), IC from the two float order statistics, PM_HD
// in double in natural order, increments computed in double and rounded to
// float.
template <class Src, class LDS>
__device__ __forceinline__ void line_setup_f32(const Src &src, const LDS &S,
                                               uint32_t &hd, uint32_t &s0, double &pm_hd,
                                               int32_t &ic, const DecodeCfg &cfg)
{
    hd = 0;
#pragma unroll 1
    for (int i = 0; i < 16; ++i) {
        const double x = src(i);
        if (!(x > 0.0)) hd |= 1u << i; // HardDec (:758)
        S.d(i) = (float)fabs(x);
    }
    s0 = syndrome16(hd);
    // ranks over the keys' bit patterns, ties by position: the keys are >= 0,
    // so this is their order (a NaN key sorts last), and always a permutation,
    // which place_by_rank's cycle walk needs
    uint64_t rank = 0;
    float a_first = 0.0f, a_eighth = 0.0f;
#pragma unroll 1
    for (int i = 0; i < 16; ++i) {
        const float ai = S.d(i);
        const uint32_t bi = __builtin_bit_cast(uint32_t, ai);
        uint32_t r = 0;
#pragma unroll 4
        for (int j = 0; j < 16; ++j) {
            const uint32_t bj = __builtin_bit_cast(uint32_t, (float)S.d(j));
            r += (bj < bi || (bj == bi && j < i)) ? 1u : 0u;
        }
        rank |= (uint64_t)r << (4 * i);
        if (r == 0) a_first = ai;
        if (r == 7) a_eighth = ai;
    }
    ic = intercept((double)a_first, (double)a_eighth, false);
    pm_hd = 0.0;
#pragma unroll 1
    for (int i = 0; i < 16; ++i) {
        const double ai = fabs(src(i));
        const double jn = jaclog(-ai, 50.0, cfg);
        pm_hd += jn;
        S.d(i) = (float)(jaclog(ai, 50.0, cfg) - jn);
    }
    place_by_rank(S, rank);
    S.d(-1) = 0.0f; // null record
    S.h(-1) = 0;
}

// Set-up in the arithmetic of P.
template <class P, class Src, class LDS>
__device__ __forceinline__ void line_setup_p(const Src &src, const LDS &S,
                                             uint32_t &hd, uint32_t &s0, double &pm_hd,
                                             int32_t &ic, const DecodeCfg &cfg)
{
    if constexpr (P::kF32)
        line_setup_f32(src, S, hd, s0, pm_hd, ic, cfg);
    else
        line_setup(src, S, hd, s0, pm_hd, ic, P::kLegacy, cfg);
}

// Record a list entry and apply the stopping rule (:628-641, :674-686, getLConf :767-774).
__device__ __forceinline__ bool list_hit(LineList &ls, double e, uint32_t code, double pnot,
                                         const DecodeCfg &cfg)
{
    ls.ppos += e;
    const double pneg = 0.00390625 * pnot; // 2^-(n-k) * P_notGuess
    ls.pnl = pneg;
    const ExpEntry et = cfg.exptab[kExpTableEntries]; // {eps, thres}
    const double gap = cfg.exptab[kLdsGapEntry].a;
    const bool confident = conf_above(ls.ppos, pneg, et.a, et.b, gap);
    ls.code |= (uint64_t)code << (16 * ls.nl);
    ls.nl += 1;
    return confident || (uint32_t)ls.nl == cfg.L;
}

// Entry q (< ls.nl) of a line's list, rebuilt from its schedule position with
// the search's own arithmetic: the same ascending path-metric sum in the
// precision of P and the same exp, so e and cw equal what the search found.
template <class P, class LDS>
__device__ __forceinline__ void list_entry(const LineList &ls, int q, const LDS &S, uint32_t hd,
                                           double pm_hd, int32_t ic, const DecodeCfg &cfg,
                                           double &e, uint32_t &cw)
{
    using Rec = typename P::Rec;
    const uint32_t code = (uint32_t)(ls.code >> (16 * q)) & 0xffffu;
    cw = hd;
    if (code == 0) {
        e = exp_neg(pm_hd, cfg.exptab);
        return;
    }
    const int tbl = ic < kSchedGenericIc ? ic : kSchedGenericIc;
    const uint32_t m = cfg.sched[(size_t)tbl * kSchedStride + (code - 1u)];
    Rec pm = (Rec)pm_hd;
#pragma unroll 1
    for (uint32_t r = m; r; r &= r - 1u) {
        const int b = __builtin_ctz(r);
        pm = pm + S.d(b);
        cw ^= 1u << (S.h(b) >> 8);
    }
    e = exp_neg((double)pm, cfg.exptab);
}

// Path metric of a pattern (mask over sorted positions) with the search's
// arithmetic: the hard decision's metric plus the increments in ascending
// sorted position (getPM, :765).
template <class P, class LDS>
__device__ __forceinline__ double pattern_pm(const LDS &S, uint32_t m, double pm_hd)
{
    using Rec = typename P::Rec;
    Rec pm = (Rec)pm_hd;
#pragma unroll 1
    for (uint32_t r = m; r; r &= r - 1u) pm = pm + S.d(__builtin_ctz(r));
    return (double)pm;
}

template <class P, class LDS>
__device__ __forceinline__ void list_entries(const LineList &ls, const LDS &S, uint32_t hd,
                                             double pm_hd, int32_t ic, const DecodeCfg &cfg,
                                             ListEntries &out)
{
#pragma unroll
    for (int q = 0; q < kMaxL; ++q) {
        out.e[q] = 0.0;
        out.cw[q] = 0;
        if (q < ls.nl) list_entry<P>(ls, q, S, hd, pm_hd, ic, cfg, out.e[q], out.cw[q]);
    }
}

// State of one line's guessing loop (sogrand_main_logic :624-744), so that
// the loop can be run in bounded steps.
struct SearchState {
    uint32_t t;       // next schedule index; the guesses so far are t + 1 (the
                      // hard decision and t patterns)
    double pnot;      // P_notGuess (:614, :626, :672, :712)
    uint32_t pm_hi;   // largest high word of the loop's path metrics (>= 0): at
                      // least 0x40800000 iff some metric reached 512
    bool done;
    const uint4 *tp;  // position-word table of the line's intercept
    uint4 nxt;        // the next chunk of 8 position words (prefetched):
    uint4 nxt2;       // first and second halves
};

// The position-word tables, at a fixed offset behind the mask tables in the
// one schedule buffer.
__device__ __forceinline__ const uint32_t *pos_tables(const uint16_t *sched)
{
    return reinterpret_cast<const uint32_t *>(reinterpret_cast<const uint8_t *>(sched) +
                                              kSchedPosOffsetBytes);
}

// Hard-decision guess (:624-641) and schedule selection.
// The hard-decision guess alone (:624-641): list, P_notGuess, done, t.
template <class P>
__device__ __forceinline__ void search_start(SearchState &ss, LineList &ls, uint32_t s0,
                                             double pm_hd, const DecodeCfg &cfg)
{
    ls.nl = 0;
    ls.pnl = 0.0;
    ls.ppos = 0.0;
    ls.code = 0;
    const double e_hd = exp_neg(pm_hd, cfg.exptab);
    if (P::kLegacy) {
        // cubic_decoder.c:590: P_notGuess stays 1; P_neg holds the last
        // guess's weight until a hit replaces it
        ss.pnot = 1.0;
        ls.pnl = e_hd;
    } else {
        ss.pnot = 1.0 - e_hd;
    }
    ss.done = false;
    if (s0 == 0) ss.done = list_hit(ls, e_hd, 0u, ss.pnot, cfg);
    ss.t = 0;
    ss.pm_hi = 0;
}

template <class P>
__device__ __forceinline__ void search_init(SearchState &ss, LineList &ls, uint32_t hd, uint32_t s0,
                                            double pm_hd, int32_t ic, const DecodeCfg &cfg)
{
    search_start<P>(ss, ls, s0, pm_hd, cfg);
    const int tbl = ic < kSchedGenericIc ? ic : kSchedGenericIc;
    ss.tp = reinterpret_cast<const uint4 *>(pos_tables(cfg.sched) + (size_t)tbl * kSchedStride);
    ss.nxt = ss.tp[0];
    ss.nxt2 = ss.tp[1];
}

// Accounting of a pair of guesses at schedule positions t and t + 1 from their
// path metrics and syndromes, strictly in schedule order (:652-744): T,
// P_notGuess, list insertion and the stopping rule; the second guess is
// dropped if the first ends the search or take2 is false. mask_of(j) gives
// the mask of guess j where a hit's metric is summed afterwards (the earlier
// decoder without a budget).
template <class P, bool kPmLoop, class LDS, class MaskOf>
__device__ __forceinline__ void pair_account(SearchState &ss, LineList &ls, const LDS &S,
                                             double pm_hd, const DecodeCfg &cfg,
                                             typename P::Rec pm1, typename P::Rec pm2,
                                             uint32_t syn1, uint32_t syn2, bool take2,
                                             uint32_t t, MaskOf mask_of)
{
    const bool hit1 = (syn1 & 0xffu) == 0, hit2 = (syn2 & 0xffu) == 0;
    double e1 = 0.0, e2 = 0.0;
    if constexpr (P::kF32) {
        // single-precision weights for the found-mass chain; a hit's list
        // weight is taken in double from its (single-precision) metric, in the
        // (rare) hit branch below
        e1 = exp_neg_f32((float)pm1);
        e2 = exp_neg_f32((float)pm2);
    } else if constexpr (kPmLoop && P::kLegacy) {
        // exact weights: the earlier decoder's P_neg is a guess's own weight
        e1 = exp_neg((double)pm1, cfg.exptab);
        e2 = exp_neg((double)pm2, cfg.exptab);
    } else if constexpr (kPmLoop) {
        e1 = exp_neg_loop((double)pm1, cfg.exptab);
        e2 = exp_neg_loop((double)pm2, cfg.exptab);
    }
    constexpr bool kStandIn = kPmLoop && !P::kLegacy && !P::kF32;
    if constexpr (kStandIn) {
        // one v_max3_u32 per pair: metrics are >= 0, so their high words order
        // like their values
        const uint32_t h1 = (uint32_t)(__builtin_bit_cast(uint64_t, (double)pm1) >> 32);
        const uint32_t h2 = (uint32_t)(__builtin_bit_cast(uint64_t, (double)pm2) >> 32);
        ss.pm_hi = max(ss.pm_hi, max(h1, h2));
    }
    ss.t = t + 1;
    if (P::kF32 && hit1) e1 = exp_neg((double)pm1, cfg.exptab);
    if (!kPmLoop && hit1) e1 = exp_neg(pattern_pm<P>(S, mask_of(0), pm_hd), cfg.exptab);
    if constexpr (!P::kLegacy)
        ss.pnot -= e1;
    else if (kPmLoop && !hit1)
        ls.pnl = e1; // cubic_decoder.c:638
    bool stop = false;
    if (hit1) {
        if constexpr (kStandIn)
            if (!((double)pm1 < 512.0)) ss.pnot = __builtin_nan("");
        stop = list_hit(ls, e1, t + 1u, ss.pnot, cfg);
    }
    // end of the schedule, or the next weight level is over budget
    if (!stop && !take2) stop = true;
    if (!stop) {
        ss.t = t + 2;
        if (P::kF32 && hit2) e2 = exp_neg((double)pm2, cfg.exptab);
        if (!kPmLoop && hit2) e2 = exp_neg(pattern_pm<P>(S, mask_of(1), pm_hd), cfg.exptab);
        if constexpr (!P::kLegacy)
            ss.pnot -= e2;
        else if (kPmLoop && !hit2)
            ls.pnl = e2;
        if (hit2) {
            if constexpr (kStandIn)
                if (!((double)pm2 < 512.0)) ss.pnot = __builtin_nan("");
            stop = list_hit(ls, e2, t + 2u, ss.pnot, cfg);
        }
    }
    ss.done = stop;
}

// a ^ b ^ c in one instruction: gfx950's three-input bit operation with the
// parity truth table 0x96 (as C, the H-column XORs of 16-bit loads come out as
// two XORs with a word select)
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c)
{
    uint32_t r;
    asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0x96" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

// Whether any lane's chunk of eight position words has a pattern with bits
// past the sixth (the words' top bit): three three-input ORs, one OR and one
// ballot per chunk instead of an OR and a ballot per pair.
__device__ __forceinline__ bool chunk_has_more(uint4 a, uint4 b)
{
    uint32_t o, p, q;
    asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0xfe" : "=v"(o) : "v"(a.x), "v"(a.y), "v"(a.z));
    asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0xfe" : "=v"(p) : "v"(a.w), "v"(b.x), "v"(b.y));
    asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0xfe" : "=v"(q) : "v"(o), "v"(p), "v"(b.z));
    return __ballot(((q | b.w) & kPosMore) != 0u) != 0ull;
}

// The four pairs of one chunk (position words a, b at schedule positions t ..
// t + 7), the second of each pair dropped once the search stops.
template <bool FINITE_T, class P, bool kMore, class LDS>
__device__ __forceinline__ void search_chunk(SearchState &ss, LineList &ls, const LDS &S,
                                             uint32_t s0, double pm_hd, int32_t ic,
                                             const DecodeCfg &cfg, uint4 a, uint4 b, uint32_t t)
{
    search_pair<FINITE_T, P, kMore>(ss, ls, S, s0, pm_hd, ic, cfg, a.x, a.y, t);
    if (!ss.done) search_pair<FINITE_T, P, kMore>(ss, ls, S, s0, pm_hd, ic, cfg, a.z, a.w, t + 2u);
    if (!ss.done) search_pair<FINITE_T, P, kMore>(ss, ls, S, s0, pm_hd, ic, cfg, b.x, b.y, t + 4u);
    if (!ss.done) search_pair<FINITE_T, P, kMore>(ss, ls, S, s0, pm_hd, ic, cfg, b.z, b.w, t + 6u);
}

// Field I of a position word (v_bfe_u32, written as the instruction: as a C
// expression the compiler folds the extract into the address scaling, one
// instruction more per field).
template <int I> __device__ __forceinline__ uint32_t pos_field(uint32_t w)
{
    uint32_t f;
    asm("v_bfe_u32 %0, %1, %2, 5" : "=v"(f) : "v"(w), "i"(1 + 5 * I));
    return f;
}

// Path metrics and syndromes of the patterns of the position words w1 and w2
// (fields 0..5: record rows, 0 = none): pm = PM_HD + the increments in
// ascending sorted position, as getPM adds them (:765), and s0 XOR the H
// columns. All twelve fields' records are read before the sums: the reads do
// not depend on the sums, so their latencies overlap (reading and summing two
// fields at a time measured 3.8 % slower in double and 10 % slower in single
// precision, even with the single-precision rounds skipped wave-uniformly where
// no lane needed them). A lane whose pattern has fewer bits reads the null
// record (row 0), which leaves both sums unchanged. Without path metrics (the
// earlier decoder without a budget) only the syndromes are formed.
template <class P, bool kPmLoop, class LDS>
__device__ __forceinline__ void pattern_sums(const LDS &S, uint32_t s0, double pm_hd, uint32_t w1,
                                             uint32_t w2, typename P::Rec &pm1,
                                             typename P::Rec &pm2, uint32_t &syn1,
                                             uint32_t &syn2)
{
    using Rec = typename P::Rec;
    const uint32_t r1[6] = {pos_field<0>(w1), pos_field<1>(w1), pos_field<2>(w1),
                            pos_field<3>(w1), pos_field<4>(w1), pos_field<5>(w1)};
    const uint32_t r2[6] = {pos_field<0>(w2), pos_field<1>(w2), pos_field<2>(w2),
                            pos_field<3>(w2), pos_field<4>(w2), pos_field<5>(w2)};
    Rec d1[6], d2[6];
    uint32_t h1[6], h2[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        if constexpr (kPmLoop) {
            d1[k] = S.dr(r1[k]);
            d2[k] = S.dr(r2[k]);
        }
        h1[k] = S.hr(r1[k]);
        h2[k] = S.hr(r2[k]);
    }
    pm1 = (Rec)pm_hd;
    pm2 = (Rec)pm_hd;
    if constexpr (kPmLoop) {
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            pm1 = pm1 + d1[k];
            pm2 = pm2 + d2[k];
        }
    }
    syn1 = xor3(xor3(xor3(s0, h1[0], h1[1]), h1[2], h1[3]), h1[4], h1[5]);
    syn2 = xor3(xor3(xor3(s0, h2[0], h2[1]), h2[2], h2[3]), h2[4], h2[5]);
}

// Two consecutive guesses of the schedule (:652-744): the patterns of the
// position words w1 and w2 at schedule positions t and t + 1. Each lane
// carries two independent path-metric and exp chains; the guesses are still
// accounted strictly in schedule order (the second is dropped if the first
// ends the search). The sums of the first six bits of each pattern come from
// pattern_sums; bits past the sixth (rare) come from the mask.
//
// Under a guess budget the guard T < Tmax runs once per weight level (:652),
// before the first pattern of each level (the words' level bit), with T = t + 1.
// In the earlier decoder (P::kLegacy) P_notGuess never changes, so a pattern's
// metric matters only if it is a codeword or, under a budget, the last guess
// (whose weight is the P_neg a budget stop leaves, cubic_decoder.c:638).
// Without a budget the loop forms syndromes alone and a hit's metric is summed
// from its mask afterwards, with the same additions.
template <bool FINITE_T, class P, bool kMore = true, class LDS>
__device__ __forceinline__ void search_pair(SearchState &ss, LineList &ls, const LDS &S,
                                            uint32_t s0, double pm_hd, int32_t ic,
                                            const DecodeCfg &cfg, uint32_t w1, uint32_t w2,
                                            uint32_t t)
{
    using Rec = typename P::Rec;
    static_assert(sizeof(Rec) == sizeof(typename LDS::Rec), "record type");
    constexpr bool kPmLoop = !(P::kLegacy && !FINITE_T);
    // the end marker is the word 0; every other word has a nonzero first field
    bool take2 = w2 != 0;
    if (FINITE_T) {
        if ((w1 & kPosLevel) && !((uint64_t)t + 1u < cfg.tmax)) {
            ss.done = true;
            return;
        }
        if (take2 && (w2 & kPosLevel) && !((uint64_t)t + 2u < cfg.tmax)) take2 = false;
    }
    if (!take2) w2 = 0;
    Rec pm1 = 0, pm2 = 0;
    uint32_t syn1, syn2;
    pattern_sums<P, kPmLoop>(S, s0, pm_hd, w1, w2, pm1, pm2, syn1, syn2);
    const int tbl = ic < kSchedGenericIc ? ic : kSchedGenericIc;
    const uint16_t *masks = cfg.sched + (size_t)tbl * kSchedStride;
    // kMore = false: no lane's chunk has a pattern with bits past the sixth
    // (search_run tests that once per chunk)
    if (kMore && __ballot(((w1 | w2) & kPosMore) != 0u) != 0ull) {
        auto more = [&](uint32_t m, Rec &pm, uint32_t &syn) {
#pragma unroll 1
            for (uint32_t i = 0; i < kPosFields; ++i) m &= m - 1u;
#pragma unroll 1
            for (; m; m &= m - 1u) {
                const int b = __builtin_ctz(m);
                syn ^= (uint32_t)S.h(b);
                if constexpr (kPmLoop) pm = pm + S.d(b);
            }
        };
        if (w1 & kPosMore) more(masks[t], pm1, syn1);
        if (w2 & kPosMore) more(masks[t + 1u], pm2, syn2);
    }
    pair_account<P, kPmLoop>(ss, ls, S, pm_hd, cfg, pm1, pm2, syn1, syn2, take2, t,
                             [&](int j) { return (uint32_t)masks[t + (uint32_t)j]; });
}

// The search of one line one guess at a time with exact weights (the same
// schedule, sums, budget guard and stopping rule as search_pair and
// pair_account, :624-744), for the rare searches whose P_notGuess chain needs
// them (chain_exact). Rolled loops: small code, few registers; it leaves the
// search loop's prefetch registers (ss.nxt, ss.nxt2), whose loads may still be
// in flight, alone.
template <bool FINITE_T, class P, class LDS>
__device__ __forceinline__ void search_exact(SearchState &ss, LineList &ls, const LDS &S,
                                             uint32_t hd, uint32_t s0, double pm_hd, int32_t ic,
                                             const DecodeCfg &cfg)
{
    search_start<P>(ss, ls, s0, pm_hd, cfg);
    const int tbl = ic < kSchedGenericIc ? ic : kSchedGenericIc;
    const uint16_t *masks = cfg.sched + (size_t)tbl * kSchedStride;
    const uint32_t *words = pos_tables(cfg.sched) + (size_t)tbl * kSchedStride;
#pragma unroll 1
    while (!ss.done) {
        const uint32_t t = ss.t;
        const uint32_t w = words[t];
        if (w == 0u || (FINITE_T && (w & kPosLevel) && !((uint64_t)t + 1u < cfg.tmax))) {
            ss.done = true; // end of the schedule, or the next weight level is over budget
            break;
        }
        double pm = pm_hd;
        uint32_t syn = s0;
#pragma unroll 1
        for (uint32_t r = masks[t]; r; r &= r - 1u) {
            const int b = __builtin_ctz(r);
            pm = pm + (double)S.d(b);
            syn ^= (uint32_t)S.h(b);
        }
        const double e = exp_neg(pm, cfg.exptab);
        ss.t = t + 1;
        ss.pnot -= e;
        if ((syn & 0xffu) == 0 && list_hit(ls, e, t + 1u, ss.pnot, cfg)) ss.done = true;
    }
}

// Further guesses of the schedule, in chunks of 8 position words (four pairs,
// 32 bytes), up to about max_steps guesses (rounded up to 16). A search starts
// at a chunk boundary and a lane that stops mid-chunk is done, so every
// searching lane enters each half-iteration at a chunk boundary; each chunk is
// loaded one chunk ahead. The loop runs an iteration count common to the wave
// and leaves when no lane is still searching; a finished lane idles inside it,
// which keeps the wave's control flow to one exec mask.
template <bool FINITE_T, class P, class LDS>
__device__ __forceinline__ void search_run(SearchState &ss, LineList &ls, const LDS &S, uint32_t hd,
                                           uint32_t s0, double pm_hd, int32_t ic,
                                           const DecodeCfg &cfg, uint32_t max_steps)
{
#pragma unroll 1
    for (uint32_t step = 0; step < max_steps; step += 16) {
        // two chunks per iteration, in alternating registers: the chunk for
        // the second half is loaded at the start of the first, the next
        // iteration's first chunk at the start of the second, each into
        // registers whose previous value is already consumed, so the load
        // lands in place and is waited for only a chunk later (with one
        // chunk per iteration the compiler copied the prefetched registers
        // after the load, which waited for it at once)
        if (__ballot(!ss.done) == 0ull) break;
        uint4 b0, b1;
        if (!ss.done) {
            const uint32_t t = ss.t; // a multiple of 8 here
            const uint4 *np = ss.tp + (t >> 2) + 2; // tables carry one chunk of padding
            b0 = np[0];
            b1 = np[1];
            // (a wave-uniform branch: the rare chunks with long patterns take
            // the variant that adds bits past the sixth; the single-precision
            // variant, held to fewer registers, tests each pair)
            if (P::kF32 || chunk_has_more(ss.nxt, ss.nxt2))
                search_chunk<FINITE_T, P, true>(ss, ls, S, s0, pm_hd, ic, cfg, ss.nxt, ss.nxt2, t);
            else
                search_chunk<FINITE_T, P, false>(ss, ls, S, s0, pm_hd, ic, cfg, ss.nxt, ss.nxt2, t);
        }
        if (__ballot(!ss.done) == 0ull) break;
        if (!ss.done) {
            const uint32_t t = ss.t;
            const uint4 *np = ss.tp + (t >> 2) + 2;
            ss.nxt = np[0];
            ss.nxt2 = np[1];
            if (P::kF32 || chunk_has_more(b0, b1))
                search_chunk<FINITE_T, P, true>(ss, ls, S, s0, pm_hd, ic, cfg, b0, b1, t);
            else
                search_chunk<FINITE_T, P, false>(ss, ls, S, s0, pm_hd, ic, cfg, b0, b1, t);
        }
    }
    if constexpr (!P::kF32 && !P::kLegacy) {
        // a search that ended with P_notGuess below 2^-680 (rounding takes it
        // to 0 or below in about 2 % of lines) and met a metric from 512 on
        // (0x40800000 is the high word of 512.0) may have needed that metric's
        // exact weight (exp_neg_loop), and one that met a NaN metric (high word
        // above +inf's 0x7ff00000; exp(-NaN) is NaN, the stand-in is not) needs
        // it: such searches are run again with exact weights
        if (ss.done && ((!chain_exact(ss.pnot) && ss.pm_hi >= 0x40800000u) ||
                        ss.pm_hi > 0x7ff00000u))
            search_exact<FINITE_T, P>(ss, ls, S, hd, s0, pm_hd, ic, cfg);
    }
}

template <bool FINITE_T, class P, class LDS>
__device__ __forceinline__ void line_search(const LDS &S, uint32_t hd,
                                            uint32_t s0, double pm_hd, int32_t ic,
                                            const DecodeCfg &cfg, LineList &ls)
{
    SearchState ss;
    search_init<P>(ss, ls, hd, s0, pm_hd, ic, cfg);
    search_run<FINITE_T, P>(ss, ls, S, hd, s0, pm_hd, ic, cfg, 0xffffffffu);
    ls.T = ss.t + 1u;
}

// Soft output of position i (SOGRAND_bitSO :536-575) given the line input x.
// cfg.eps: EPSILON of the clamps (1e-40, :16; DBL_EPSILON in
// cubic_decoder_updated.c:16; 1e-9 in cubic_decoder.c:509-510); cfg.eps_log:
// floor inside the logs (EPSILON, or 1e-30 in cubic_decoder.c:533).
__device__ __forceinline__ double soft_app(const ListEntries &ls, int nl, double pnotin, int i,
                                           double x, const DecodeCfg &cfg)
{
    // eps from the LDS table entry, read here: 1 - eps is then formed where it
    // is used rather than kept live (and spilled) for the whole kernel
    const double eps = cfg.exptab[kExpTableEntries].a, eps_log = cfg.eps_log;
    double pp1 = 1.0 / (1.0 + exp_any(x, cfg.exptab));
    double pp0 = 1.0 - pp1;
    pp1 = fmin(fmax(pp1, eps), 1.0 - eps);
    pp0 = fmin(fmax(pp0, eps), 1.0 - eps);
    double p1 = 0.0, p0 = 0.0;
#pragma unroll
    for (int j = 0; j < kMaxL; ++j)
        if (j < nl) {
            double c = (double)((ls.cw[j] >> i) & 1u);
            p1 += c * ls.e[j];
            p0 += (1.0 - c) * ls.e[j];
        }
    p0 += pnotin * pp0;
    p1 += pnotin * pp1;
    return log_pos(fmax(p0, eps_log), cfg.logtab, log_coefs(cfg)) -
           log_pos(fmax(p1, eps_log), cfg.logtab, log_coefs(cfg));
}

} // namespace sg
