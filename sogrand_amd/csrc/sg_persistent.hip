// Persistent batched decoder for gfx950.
//
// Each wave is an independent worker that keeps S cubes in flight ("slots").
// A slot holds one cube's current phase; its 256 lines are handed to lanes as
// lanes become free, so a lane that finishes a short line immediately takes
// another line, of the same cube or of another slot's cube, instead of idling
// until the slowest line of the phase is done. When a phase's 256 lines are
// in, the wave runs the early-termination test on the slot's hard-decision
// bitmap and either starts the next phase or retires the cube; a retired slot
// pops the next cube index from a device-wide counter. The kernel returns when
// the counter is exhausted and every slot is empty, so every wave reaches the
// same exit.
//
// Per-line arithmetic is the shared line decoder (sg_device.h), so results are
// bit-identical to the plain kernel and independent of batch composition.
// Semantics: SOGRAND_C/cubic_decoder_final.c:119-289 (per-codeword loop).
#include <hip/hip_runtime.h>

#include "sg_common.h"
#include "sg_device.h"
#include "sg_launch.h"

#pragma clang fp contract(off)

namespace sg {

namespace {

constexpr int kWaves = 4; // waves per workgroup (independent)

struct CwStatsOut {
    uint32_t half_iters;
    uint32_t converged;
    uint64_t NG;
    uint64_t NG_p;
};

// Device scratch per cube slot, in cubes of 4096 doubles: two L_E cubes used in
// turn, then the inputs x = L_ch + alpha L_E of the current phase's lines.
// The set-up stores x line-major (16 line + i: a line's 16 values are one
// 128-byte row), where the soft output of the same line reads it back. The L_E
// of a phase goes to one L_E cube and the next phase reads it from there while
// writing the other (phase_geom gives both layouts).
constexpr int kSlotScratch = kPersistentSlotScratch;
static_assert(kSlotScratch == 3, "two L_E cubes and the line inputs");

// Where the lines of a phase sit. Every 16 lanes of an instruction work on
// the 16 positions of one line, four lines an instruction, and lines finish in
// no particular order. Loads want four consecutive lines to read 32 contiguous
// bytes at each position; stores want a finished line's 16 values in one
// 128-byte row. Per axis (a0 + 16 a1 + 256 a2 natural, lo = line & 15,
// hi = line >> 4):
//   axis 0: lines a2 = lo, a1 = hi (natural rows); reads axis 2's rows at
//           lo + 256 hi + 16 i; writes its rows, 16 line + i
//   axis 1: lines a0 = lo, a2 = hi; reads axis 0's rows at lo + 16 hi + 256 i;
//           writes natural order (single doubles 128 bytes apart)
//   axis 2: lines a0 = lo, a1 = hi; reads natural order; writes its rows
// The channel LLRs are read in natural order: whole rows on axis 0, 32 bytes
// per four lines on axes 1 and 2. Only axis 1's stores are scattered: writing
// rows there would leave axis 2 either the L_E or the channel loads at single
// doubles 128 bytes apart, which the counters show to cost more.
struct PhaseGeom {
    int base, stride;   // natural positions base + stride i
    int rbase, rstride; // L_E read: rbase + rstride i in the previous phase's cube
    int wbase, wstride; // L_E write: wbase + wstride i in this phase's cube
};
__device__ __forceinline__ PhaseGeom phase_geom(int ax, int line)
{
    const int lo = line & 15, hi = line >> 4;
    PhaseGeom g;
    if (ax == 0) {
        g.base = 256 * lo + 16 * hi;
        g.stride = 1;
        g.rbase = lo + 256 * hi;
        g.rstride = 16;
        g.wbase = 16 * line;
        g.wstride = 1;
    } else if (ax == 1) {
        g.base = lo + 256 * hi;
        g.stride = 16;
        g.rbase = lo + 16 * hi;
        g.rstride = 256;
        g.wbase = g.base;
        g.wstride = g.stride;
    } else {
        g.base = lo + 16 * hi;
        g.stride = 256;
        g.rbase = g.base;
        g.rstride = g.stride;
        g.wbase = 16 * line;
        g.wstride = 1;
    }
    return g;
}

// Early-termination test of one 16x16 bit plane streamed word by word: word j
// (bit q = column q) is checked as a component codeword itself when `rows` is
// set, and folded into the 8 column syndromes X_r = W[8+r] ^ XOR_{j<8, P[j][r]}
// W[j], which vanish iff every column q is a component codeword.
template <class Word>
__device__ __forceinline__ bool plane_ok(const Word &word, bool rows)
{
    uint32_t x[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    bool ok = true;
#pragma unroll 1
    for (int j = 0; j < 16; ++j) {
        const uint32_t w = word(j);
        if (rows) ok = ok && (syndrome16(w) == 0);
        const uint32_t h = hcol_of((unsigned)j); // column j of H: rows r it feeds
#pragma unroll
        for (int r = 0; r < 8; ++r)
            if ((h >> r) & 1u) x[r] ^= w;
    }
    uint32_t bad = 0;
#pragma unroll
    for (int r = 0; r < 8; ++r) bad |= x[r];
    return ok && (bad & 0xffffu) == 0;
}

} // namespace

// Lane states of the persistent kernel.
enum : int { kIdle = 0, kSearch = 1, kDone = 2, kNew = 3 };

// Wave-uniform state of a wave's cube slots. The arrays are only ever indexed
// by constants (fully unrolled slot loops); what a lane needs for the slot of
// its own line (cube, axis, iteration) is mirrored in LDS, where a per-lane
// index is cheap.
template <int S> struct SlotState {
    uint32_t cube[S], it[S], ax[S], nxt_line[S], half[S];
    uint64_t ng[S], ngp[S];
    bool active[S];
};

// The lowest four set lanes of a wave mask (owners of one transition sub-round),
// -1 where there are fewer.
__device__ __forceinline__ void take4(uint64_t &m, int (&o)[4])
{
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (m) {
            o[q] = __builtin_ctzll(m);
            m &= m - 1ull;
        } else {
            o[q] = -1;
        }
    }
}

__device__ __forceinline__ double shfl_d(double v, int src)
{
    const uint64_t u = double_to_bits(v);
    const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)u, src, 64);
    const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(u >> 32), src, 64);
    return bits_to_double(((uint64_t)hi << 32) | lo);
}

// XOR of v over each group of 16 lanes (lanes 16q .. 16q + 15), in every lane
// of the group: two quad permutations, then the half-row and row mirrors.
__device__ __forceinline__ uint32_t xor16(uint32_t v)
{
    v ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, true);  // quad_perm 1,0,3,2
    v ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, true);  // quad_perm 2,3,0,1
    v ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xF, 0xF, true); // row_half_mirror
    v ^= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x140, 0xF, 0xF, true); // row_mirror
    return v;
}

// OR of v over each group of 16 lanes, in every lane of the group (the
// pattern of xor16).
__device__ __forceinline__ uint32_t or16(uint32_t v)
{
    v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, true);
    v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, true);
    v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xF, 0xF, true);
    v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x140, 0xF, 0xF, true);
    return v;
}

// Four waves per SIMD asked for: the single-precision variant's records fit
// four workgroups' LDS per CU and its registers fit 128 with a few bytes of
// spill, which measures faster than three waves; the FP64 kernel's LDS allows
// three workgroups per CU whatever the hint says (and asking for three
// instead measures 17 % slower: the search loop is scheduled differently).
template <bool FINITE_T, int S, class P>
__global__ __launch_bounds__(64 * kWaves) __attribute__((amdgpu_waves_per_eu(4))) void persistent_kernel(CubeJob job, uint32_t thresh,
                                                                 uint32_t step_len)
{
    using Rec = typename P::Rec;
    __shared__ Rec s_d[kWaves][17 * 64];
    __shared__ uint16_t s_h[kWaves][17 * 64];
    __shared__ uint32_t s_hd[kWaves][S][128];
    __shared__ uint32_t s_cnt[kWaves][S][3]; // lines done, phase guess sum, phase max
    __shared__ double s_stage_jn[kWaves][64];
    __shared__ double s_stage_a[kWaves][4][2]; // per sub-round line: |x| of ranks 0 and 7
    __shared__ uint32_t s_geom[kWaves][S][2];  // per slot: cube, axis | iteration << 8
    __shared__ ExpEntry s_exp[kLdsTableEntries];

    load_exp_table(s_exp, threadIdx.x, job.eps, job.thres);
    __syncthreads(); // the only workgroup barrier: waves run independently from here
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    Rec *rec_d = &s_d[wave][0];
    uint16_t *rec_h = &s_h[wave][0];
    const LineLdsT<Rec, 64> L(rec_d, rec_h, lane);
    const DecodeCfg cfg{job.tmax, job.thres, job.sched, job.L, s_exp, kLogTable, job.eps, job.eps_log};
    const size_t wave_id = (size_t)blockIdx.x * kWaves + wave;
    uint32_t *hd_bits = &s_hd[wave][0][0];
    uint32_t *cnt = &s_cnt[wave][0][0];
    double *stage_jn = &s_stage_jn[wave][0];
    uint32_t *geom = &s_geom[wave][0][0];
    double *stage_a = &s_stage_a[wave][0][0];

    SlotState<S> sl;
#pragma clang loop unroll(full)
    for (int s = 0; s < S; ++s) {
        sl.active[s] = false;
        sl.cube[s] = sl.it[s] = sl.ax[s] = sl.nxt_line[s] = sl.half[s] = 0;
        sl.ng[s] = sl.ngp[s] = 0;
    }
    bool queue_empty = false;

    // lane state: the line this lane owns and its search
    int state = kIdle, my_slot = 0, my_line = 0;
    uint32_t hd = 0, s0 = 0;
    double pm_hd = 0.0;
    int32_t ic = 0;
    LineList ls;
    SearchState ss;
    ls.nl = 0;
    ls.pnl = 0.0;
    ss.t = ~0u; // no line yet: t + 1 = 0 guesses
    ss.done = true;

    // geometry of (slot, line) as seen by any lane: positions (phase_geom),
    // the L_E cubes this phase writes (le_wr) and reads (le_rd), the line
    // inputs (xs, line-major)
    struct LineGeom {
        PhaseGeom g;
        const double *lch;
        const double *le_rd;
        double *le_wr, *xs, *soft;
        bool first;
    };
    auto line_geom = [&](int s, int line) {
        LineGeom lg;
        const uint32_t c = geom[2 * s], g = geom[2 * s + 1];
        const uint32_t ax = g & 0xffu, it = g >> 8;
        lg.g = phase_geom((int)ax, line);
        lg.lch = job.llr + (size_t)c * kCube;
        double *sc = job.le_scratch + (wave_id * S + (size_t)s) * (kSlotScratch * kCube);
        // phases alternate the two L_E cubes: phase 3 (it - 1) + ax + 1 writes
        // cube (it + ax) & 1 and reads the other one
        const uint32_t wb = (it + ax) & 1u;
        lg.le_wr = sc + (size_t)wb * kCube;
        lg.le_rd = sc + (size_t)(wb ^ 1u) * kCube;
        lg.xs = sc + 2 * kCube;
        lg.soft = job.soft ? job.soft + (size_t)c * kCube : nullptr;
        lg.first = (it == 1 && ax == 0);
        return lg;
    };

    for (;;) {
        // ---- 1. phase completion / cube retirement (wave-uniform) ----
#pragma clang loop unroll(full)
        for (int s = 0; s < S; ++s) {
            if (!sl.active[s] || sl.nxt_line[s] < 256) continue;
            const uint32_t done_lines = __builtin_amdgcn_readfirstlane(cnt[3 * s + 0]);
            if (done_lines < 256) continue;
            // all lines of the phase are in: make their L_E stores visible to
            // the next phase's loads
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
            sl.ng[s] += __builtin_amdgcn_readfirstlane(cnt[3 * s + 1]);
            sl.ngp[s] += __builtin_amdgcn_readfirstlane(cnt[3 * s + 2]);
            // early termination: all 768 lines of the hard decision are
            // component codewords (equivalent to the re-encode test :317-391)
            bool ok = true;
            const uint32_t *hb = hd_bits + 128 * s;
            if (lane < 32) {
                // lanes 0-15: plane a2 = q (words over a1: axis-0 lines, columns =
                // axis-1 lines); lanes 16-31: plane a1 = q (columns = axis-2 lines)
                const int q = lane & 15;
                const bool a2_plane = lane < 16;
                auto word = [&](int j) {
                    const int h = a2_plane ? (j + 16 * q) : (q + 16 * j);
                    return (hb[h >> 1] >> (16 * (h & 1))) & 0xffffu;
                };
                ok = plane_ok(word, a2_plane);
            }
            const bool conv = __ballot(!ok) == 0ull;
            const bool last = (sl.ax[s] == 2 && sl.it[s] >= job.Imax);
            if (conv || last) {
                // retire: info bits from the hard decision (:267-280), statistics
                const int h = (lane & 7) + 16 * (lane >> 3);
                const uint32_t byte = rev8((hb[h >> 1] >> (16 * (h & 1))) & 0xffu);
                // the lane offset is made opaque here so that the compiler does
                // not keep job.info + lane live (in spilled registers) for the
                // whole kernel
                int lo = lane;
                asm volatile("" : "+v"(lo));
                job.info[(size_t)sl.cube[s] * kInfoBytes + (size_t)lo] = (uint8_t)byte;
                if (lane == 0) {
                    if (job.stats) {
                        CwStatsOut *o = reinterpret_cast<CwStatsOut *>(job.stats) + sl.cube[s];
                        o->half_iters = sl.half[s];
                        o->converged = conv ? 1u : 0u;
                        o->NG = sl.ng[s];
                        o->NG_p = sl.ngp[s];
                    }
                }
                sl.active[s] = false;
            } else {
                if (sl.ax[s] == 2) {
                    sl.ax[s] = 0;
                    sl.it[s] += 1;
                } else {
                    sl.ax[s] += 1;
                }
                sl.half[s] += 1; // n_iter += 0.5 before each phase (:145, :184, :223)
                if (lane == 0) geom[2 * s + 1] = sl.ax[s] | (sl.it[s] << 8);
                sl.nxt_line[s] = 0;
                hd_bits[128 * s + lane] = 0;
                hd_bits[128 * s + 64 + lane] = 0;
                if (lane < 3) cnt[3 * s + lane] = 0;
            }
        }
        // ---- 2. refill empty slots from the device queue ----
#pragma clang loop unroll(full)
        for (int s = 0; s < S; ++s) {
            if (sl.active[s] || queue_empty) continue;
            uint32_t c = 0;
            if (lane == 0) c = atomicAdd(job.queue, 1u);
            c = __builtin_amdgcn_readfirstlane(c);
            if (c >= job.n_cw) {
                queue_empty = true;
                continue;
            }
            sl.active[s] = true;
            sl.cube[s] = c;
            sl.it[s] = 1;
            sl.ax[s] = 0;
            sl.half[s] = 1;
            sl.nxt_line[s] = 0;
            sl.ng[s] = sl.ngp[s] = 0;
            if (lane == 0) {
                geom[2 * s] = c;
                geom[2 * s + 1] = 0u | (1u << 8);
            }
            hd_bits[128 * s + lane] = 0;
            hd_bits[128 * s + 64 + lane] = 0;
            if (lane < 3) cnt[3 * s + lane] = 0;
        }
        bool any_active = false;
        uint32_t avail = 0;
#pragma clang loop unroll(full)
        for (int s = 0; s < S; ++s) {
            any_active = any_active || sl.active[s];
            if (sl.active[s]) avail += 256u - sl.nxt_line[s];
        }
        if (!any_active) break;

        // ---- 3. transitions, cooperatively, four lines per sub-round ----
        const uint32_t n_done = (uint32_t)__popcll(__ballot(state == kDone));
        const uint32_t n_idle = (uint32_t)__popcll(__ballot(state == kIdle));
        const uint32_t n_search = (uint32_t)__popcll(__ballot(state == kSearch));
        const uint32_t n_new = n_idle < avail ? n_idle : avail;
        if (n_done + n_new >= thresh || n_search == 0) {
            // 3a. soft output of finished lines (SOGRAND_bitSO :536-575): lane
            // 16q + i computes position i of the q-th finished line
            uint64_t D = __ballot(state == kDone);
            // the line inputs stored by the set-up (possibly in this same
            // round's previous pass, for a line that stopped at once) are
            // complete before they are read back
            if (D) __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
            while (D) {
                int o[4];
                take4(D, o);
                const int q = lane >> 4, i = lane & 15;
                const int owner = q == 0 ? o[0] : q == 1 ? o[1] : q == 2 ? o[2] : o[3];
                const int src = owner >= 0 ? owner : 0;
                const int o_slot = __shfl(my_slot, src, 64);
                const int o_line = __shfl(my_line, src, 64);
                LineList ol;
                ol.nl = __shfl(ls.nl, src, 64);
                ol.pnl = shfl_d(ls.pnl, src);
                const uint32_t code_lo = (uint32_t)__shfl((int)(uint32_t)ls.code, src, 64);
                const uint32_t code_hi = (uint32_t)__shfl((int)(uint32_t)(ls.code >> 32), src, 64);
                ol.code = ((uint64_t)code_hi << 32) | code_lo;
                // lanes 16q + j (j < list size) rebuild entry j of the owner's
                // list from its records, then the entries go to all 16 lanes
                double my_e = 0.0;
                uint32_t my_cw = 0;
                {
                    const uint32_t o_hd = (uint32_t)__shfl((int)hd, src, 64);
                    const int32_t o_ic = __shfl(ic, src, 64);
                    const double o_pm_hd = shfl_d(pm_hd, src);
                    if (owner >= 0 && i < ol.nl) {
                        const LineLdsT<Rec, 64> R(rec_d, rec_h, src);
                        list_entry<P>(ol, i, R, o_hd, o_pm_hd, o_ic, cfg, my_e, my_cw);
                    }
                }
                ListEntries ent;
#pragma unroll
                for (int j = 0; j < kMaxL; ++j) {
                    ent.e[j] = shfl_d(my_e, (lane & ~15) + j);
                    ent.cw[j] = (uint32_t)__shfl((int)my_cw, (lane & ~15) + j, 64);
                }
                if (owner >= 0) {
                    const LineGeom lg = line_geom(o_slot, o_line);
                    const int p = lg.g.base + i * lg.g.stride;
                    const double xin = lg.xs[16 * o_line + i]; // the set-up's x
                    double la, lenew;
                    if (ol.nl == 0) { // empty list: L_APP = llr, L_E = 0 (:537-539)
                        la = xin;
                        lenew = 0.0;
                    } else {
                        la = soft_app(ent, ol.nl, fmax(ol.pnl, cfg.eps), i, xin, cfg);
                        lenew = la - xin;
                    }
                    lg.le_wr[lg.g.wbase + i * lg.g.wstride] = lenew;
                    if (lg.soft) lg.soft[p] = la;
                    if (!(la > 0.0)) atomicOr(&hd_bits[128 * o_slot + (p >> 5)], 1u << (p & 31));
                }
                if (lane == o[0] || lane == o[1] || lane == o[2] || lane == o[3]) {
                    atomicAdd(&cnt[3 * my_slot + 1], ss.t + 1u);
                    atomicMax(&cnt[3 * my_slot + 2], ss.t + 1u);
                    atomicAdd(&cnt[3 * my_slot + 0], 1u);
                    state = kIdle;
                }
            }
            // 3b. hand the next lines of the active phases to free lanes
            uint64_t idle = __ballot(state == kIdle);
#pragma clang loop unroll(full)
            for (int s = 0; s < S; ++s) {
                if (!sl.active[s] || sl.nxt_line[s] >= 256 || idle == 0ull) continue;
                const uint32_t left = 256u - sl.nxt_line[s];
                const uint32_t ni = (uint32_t)__popcll(idle);
                const uint32_t k = ni < left ? ni : left;
                // idle lanes below this one (v_mbcnt: no lane mask kept live)
                const uint32_t rank = __builtin_amdgcn_mbcnt_hi(
                    (uint32_t)(idle >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
                if (state == kIdle && rank < k) {
                    state = kNew;
                    my_slot = s;
                    my_line = (int)(sl.nxt_line[s] + rank);
                }
                sl.nxt_line[s] += k;
                idle = __ballot(state == kIdle);
            }
            // 3c. set-up of the new lines (sogrand_main_logic :598-622)
            uint64_t N = __ballot(state == kNew);
            while (N) {
                int o[4];
                take4(N, o);
                const int q = lane >> 4, i = lane & 15;
                const int owner = q == 0 ? o[0] : q == 1 ? o[1] : q == 2 ? o[2] : o[3];
                const int src = owner >= 0 ? owner : 0;
                const int o_slot = __shfl(my_slot, src, 64);
                const int o_line = __shfl(my_line, src, 64);
                const LineLdsT<Rec, 64> R(rec_d, rec_h, src); // the owner's records
                bool tie = false, neg = false;
                double ai = 0.0, delta = 0.0;
                Rec key = 0;
                uint32_t rank = 0;
                if (owner >= 0) {
                    const LineGeom lg = line_geom(o_slot, o_line);
                    const int p = lg.g.base + i * lg.g.stride;
                    const double e = lg.first ? 0.0 : lg.le_rd[lg.g.rbase + i * lg.g.rstride];
                    const double tt = job.alpha * e; // L_A = alpha L_E, then + L_channel (:147-155)
                    const double x = lg.lch[p] + tt;
                    lg.xs[16 * o_line + i] = x; // read back by the soft output
                    neg = !(x > 0.0); // HardDec (:758)
                    ai = fabs(x);
                    key = (Rec)ai;
                    R.d(i) = key;
                }
                if (owner >= 0) {
                    // rank by counting. In double, the number of smaller
                    // magnitudes, a permutation of 0..15 over the line's lanes
                    // exactly when its magnitudes are distinct; a line with
                    // equal magnitudes (a tie) or a NaN goes to the owner, which
                    // replays the reference's quicksort. In the single-precision
                    // variant the keys' bit patterns are ranked, equal keys by
                    // position: a permutation whatever the keys (NaN last).
                    if constexpr (P::kF32) {
                        const uint32_t bk = __builtin_bit_cast(uint32_t, (float)key);
#pragma unroll 4
                        for (int j = 0; j < 16; ++j) {
                            const uint32_t bj = __builtin_bit_cast(uint32_t, (float)R.d(j));
                            rank += (bj < bk || (bj == bk && j < i)) ? 1u : 0u;
                        }
                    } else {
#pragma unroll 4
                        for (int j = 0; j < 16; ++j) rank += R.d(j) < key ? 1u : 0u;
                    }
                    const double jn = jaclog(-ai, P::kJacLim, cfg);
                    delta = jaclog(ai, P::kJacLim, cfg) - jn; // getPM's increment (:765)
                    stage_jn[lane] = jn;
                    // the two order statistics of IC (a tie line's IC is taken
                    // from its sorted records instead)
                    if (rank == 0) stage_a[2 * q] = (double)key;
                    if (rank == 7) stage_a[2 * q + 1] = (double)key;
                }
                // a repeated rank leaves one of the line's 16 rank bits unset
                if constexpr (!P::kF32) tie = or16(owner >= 0 ? 1u << rank : 0u) != 0xffffu;
                const uint64_t special = __ballot(owner >= 0 && tie);
                const uint64_t negm = __ballot(owner >= 0 && neg);
                // syndrome of the hard decision: XOR of the H columns of the
                // line's negative positions over its 16 lanes
                const uint32_t syn16 = xor16(neg ? (uint32_t)hcol_of((unsigned)i) : 0u);
                // owners: hard decision, PM_HD in natural order, intercept
                const int myq = lane == o[0] ? 0 : lane == o[1] ? 1 : lane == o[2] ? 2
                              : lane == o[3] ? 3 : -1;
                // the owner of line q takes the syndrome from the line's lanes
                const uint32_t syn_owner = (uint32_t)__shfl((int)syn16, 16 * (myq >= 0 ? myq : 0), 64);
                bool line_special = false;
                if (myq >= 0) {
                    line_special = ((special >> (16 * myq)) & 0xffffull) != 0;
                    hd = (uint32_t)(negm >> (16 * myq)) & 0xffffu;
                    s0 = syn_owner;
                    pm_hd = 0.0;
#pragma unroll
                    for (int j = 0; j < 16; ++j) pm_hd += stage_jn[16 * myq + j]; // getPM_HD (:617)
                    // IC (:619-622); a tie line's is taken from its sorted records
                    ic = intercept(stage_a[2 * myq], stage_a[2 * myq + 1], P::kLegacy);
                }
                // tasks of ordinary lines write the sorted records
                if (owner >= 0 && ((special >> (16 * q)) & 0xffffull) == 0) {
                    R.d((int)rank) = (Rec)delta;
                    R.h((int)rank) = (uint16_t)(hcol_of((unsigned)i) | (i << 8));
                }
                if (myq >= 0) {
                    L.d(-1) = (Rec)0; // null record read by empty position-word fields
                    L.h(-1) = 0;
                    if constexpr (!P::kF32) {
                        // records hold |x| in natural order
                        if (line_special) {
                            double a_first, a_eighth;
                            pm_hd = setup_with_ties(L, P::kJacLim, cfg, a_first, a_eighth);
                            ic = intercept(a_first, a_eighth, P::kLegacy);
                        }
                    }
                    search_init<P>(ss, ls, hd, s0, pm_hd, ic, cfg);
                    state = ss.done ? kDone : kSearch;
                }
            }
        }
        // ---- 4. a bounded step of every searching lane's guessing loop ----
        if (state == kSearch) {
            search_run<FINITE_T, P>(ss, ls, L, hd, s0, pm_hd, ic, cfg, step_len);
            if (ss.done) state = kDone;
        }
    }
}

int persistent_blocks_per_cu(int slots, bool finite, int arith)
{
    int nb = 0;
    hipError_t e = hipErrorInvalidValue;
#define SG_OCC(FT, SS, PP)                                                                    \
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, persistent_kernel<FT, SS, PP>,       \
                                                     64 * kWaves, 0)
#define SG_OCC_P(FT, SS)                                                                      \
    if (arith == kArithF32) SG_OCC(FT, SS, PrecF32);                                          \
    else if (arith == kArithF64Legacy) SG_OCC(FT, SS, PrecF64Legacy);                         \
    else SG_OCC(FT, SS, PrecF64)
    if (slots == 1) {
        if (finite) { SG_OCC_P(true, 1); } else { SG_OCC_P(false, 1); }
    } else if (slots == 2) {
        if (finite) { SG_OCC_P(true, 2); } else { SG_OCC_P(false, 2); }
    } else if (slots == 3) {
        if (finite) { SG_OCC_P(true, 3); } else { SG_OCC_P(false, 3); }
    } else if (slots == 4) {
        if (finite) { SG_OCC_P(true, 4); } else { SG_OCC_P(false, 4); }
    }
#undef SG_OCC_P
#undef SG_OCC
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return 1;
    }
    return nb > 0 ? nb : 1;
}

hipError_t launch_persistent(const CubeJob &job, int waves_per_block, int blocks, int slots,
                             uint32_t lane_batch, uint32_t step_len, hipStream_t st)
{
    if (job.n_cw == 0) return hipSuccess;
    if (waves_per_block != kWaves) return hipErrorInvalidValue;
    const bool finite = tmax_binds(job.tmax);
    const uint32_t thresh = lane_batch, step = step_len;
    const dim3 g(blocks), b(64 * kWaves);
#define SG_LAUNCH(FT, SS, PP)                                                                 \
    hipLaunchKernelGGL((persistent_kernel<FT, SS, PP>), g, b, 0, st, job, thresh, step)
#define SG_LAUNCH_P(FT, SS)                                                                   \
    if (job.arith == kArithF32) SG_LAUNCH(FT, SS, PrecF32);                                   \
    else if (job.arith == kArithF64Legacy) SG_LAUNCH(FT, SS, PrecF64Legacy);                  \
    else SG_LAUNCH(FT, SS, PrecF64)
    if (slots == 1) {
        if (finite) { SG_LAUNCH_P(true, 1); } else { SG_LAUNCH_P(false, 1); }
    } else if (slots == 2) {
        if (finite) { SG_LAUNCH_P(true, 2); } else { SG_LAUNCH_P(false, 2); }
    } else if (slots == 3) {
        if (finite) { SG_LAUNCH_P(true, 3); } else { SG_LAUNCH_P(false, 3); }
    } else if (slots == 4) {
        if (finite) { SG_LAUNCH_P(true, 4); } else { SG_LAUNCH_P(false, 4); }
    } else {
        return hipErrorInvalidValue;
    }
#undef SG_LAUNCH_P
#undef SG_LAUNCH
    return hipGetLastError();
}

} // namespace sg
