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

__device__ __forceinline__ void axis_geom(int ax, int line, int &base, int &stride)
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

struct PSrc {
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

// Bit-sliced component check of the 16 columns of a 16x16 bit plane W
// (bit q of W[j] = row j, column q): column q is a codeword iff all 8 syndrome
// rows vanish; X_r = W[8+r] ^ XOR_{j<8, P[j][r]=1} W[j].
__device__ __forceinline__ bool plane_cols_ok(const uint32_t (&W)[16])
{
    uint32_t bad = 0;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        uint32_t x = W[8 + r];
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if ((hcol(j) >> r) & 1u) x ^= W[j];
        bad |= x;
    }
    return (bad & 0xffffu) == 0;
}

} // namespace

template <bool FINITE_T, int S>
__global__ __launch_bounds__(64 * kWaves) void persistent_kernel(CubeJob job, uint32_t thresh,
                                                                 uint32_t step_len)
{
    __shared__ double s_d[kWaves][16 * 64];
    __shared__ uint16_t s_h[kWaves][16 * 64];
    __shared__ uint32_t s_hd[kWaves][S][128];
    __shared__ uint32_t s_cnt[kWaves][S][4]; // lines done, phase guess sum, phase max, bad

    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint64_t lt_mask = (1ull << lane) - 1ull;
    const LineLds L{&s_d[wave][0], &s_h[wave][0], 64, lane};
    const DecodeCfg cfg{job.tmax, job.thres, job.sched, job.L};
    const size_t wave_id = (size_t)blockIdx.x * kWaves + wave;
    uint32_t *hd_bits = &s_hd[wave][0][0];
    uint32_t *cnt = &s_cnt[wave][0][0];

    // wave-uniform slot state
    uint32_t cube[S], it[S], ax[S], nxt_line[S], half[S];
    uint64_t ng[S], ngp[S];
    bool active[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        active[s] = false;
        cube[s] = it[s] = ax[s] = nxt_line[s] = half[s] = 0;
        ng[s] = ngp[s] = 0;
    }
    bool queue_empty = false;

    // lane state
    int my_slot = -1;
    const double *lch = nullptr;
    double *le = nullptr, *soft = nullptr;
    int base = 0, stride = 1;
    bool first = false;
    uint32_t hd = 0, s0 = 0;
    double pm_hd = 0.0;
    int32_t ic = 0;
    LineList ls;
    SearchState ss;
    ss.done = true;

    for (;;) {
        // ---- 1. phase completion / cube retirement (wave-uniform) ----
#pragma unroll
        for (int s = 0; s < S; ++s) {
            if (!active[s] || nxt_line[s] < 256) continue;
            const uint32_t done_lines = __builtin_amdgcn_readfirstlane(cnt[4 * s + 0]);
            if (done_lines < 256) continue;
            // all lines of the phase are in: make their L_E stores visible to
            // the next phase's loads
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
            ng[s] += __builtin_amdgcn_readfirstlane(cnt[4 * s + 1]);
            ngp[s] += __builtin_amdgcn_readfirstlane(cnt[4 * s + 2]);
            // early termination: all 768 lines of the hard decision are
            // component codewords (equivalent to the re-encode test :317-391)
            bool ok = true;
            const uint32_t *hb = hd_bits + 128 * s;
            if (lane < 32) {
                uint32_t W[16];
                const int q = lane & 15;
#pragma unroll
                for (int j = 0; j < 16; ++j) {
                    // lanes 0-15: plane a2 = q, words over a1; lanes 16-31: plane a1 = q, words over a2
                    const int h = lane < 16 ? (j + 16 * q) : (q + 16 * j);
                    W[j] = (hb[h >> 1] >> (16 * (h & 1))) & 0xffffu;
                }
                ok = plane_cols_ok(W);
                if (lane < 16) {
#pragma unroll
                    for (int j = 0; j < 16; ++j) ok = ok && (syndrome16(W[j]) == 0);
                }
            }
            const bool conv = __ballot(!ok) == 0ull;
            const bool last = (ax[s] == 2 && it[s] >= job.Imax);
            if (conv || last) {
                // retire: info bits from the hard decision (:267-280), statistics
                const int h = (lane & 7) + 16 * (lane >> 3);
                const uint32_t byte = rev8((hb[h >> 1] >> (16 * (h & 1))) & 0xffu);
                job.info[(size_t)cube[s] * kInfoBytes + lane] = (uint8_t)byte;
                if (lane == 0) {
                    if (job.stats) {
                        CwStatsOut *o = reinterpret_cast<CwStatsOut *>(job.stats) + cube[s];
                        o->half_iters = half[s];
                        o->converged = conv ? 1u : 0u;
                        o->NG = ng[s];
                        o->NG_p = ngp[s];
                    }
                    job.status[cube[s]] = cnt[4 * s + 3] ? 1 : 0;
                }
                active[s] = false;
            } else {
                if (ax[s] == 2) {
                    ax[s] = 0;
                    it[s] += 1;
                } else {
                    ax[s] += 1;
                }
                half[s] += 1; // n_iter += 0.5 before each phase (:145, :184, :223)
                nxt_line[s] = 0;
                hd_bits[128 * s + lane] = 0;
                hd_bits[128 * s + 64 + lane] = 0;
                if (lane < 3) cnt[4 * s + lane] = 0;
            }
        }
        // ---- 2. refill empty slots from the device queue ----
#pragma unroll
        for (int s = 0; s < S; ++s) {
            if (active[s] || queue_empty) continue;
            uint32_t c = 0;
            if (lane == 0) c = atomicAdd(job.queue, 1u);
            c = __builtin_amdgcn_readfirstlane(c);
            if (c >= job.n_cw) {
                queue_empty = true;
                continue;
            }
            active[s] = true;
            cube[s] = c;
            it[s] = 1;
            ax[s] = 0;
            half[s] = 1;
            nxt_line[s] = 0;
            ng[s] = ngp[s] = 0;
            hd_bits[128 * s + lane] = 0;
            hd_bits[128 * s + 64 + lane] = 0;
            if (lane < 4) cnt[4 * s + lane] = 0;
        }
        bool any_active = false;
#pragma unroll
        for (int s = 0; s < S; ++s) any_active = any_active || active[s];
        if (!any_active) break;

        // ---- 3. hand lines to free lanes, in batches ----
        uint64_t idle = __ballot(my_slot < 0);
        const uint32_t n_idle = (uint32_t)__popcll(idle);
        bool fresh = false;
        if (n_idle >= thresh || n_idle == 64u) {
#pragma unroll
            for (int s = 0; s < S; ++s) {
                if (!active[s] || nxt_line[s] >= 256 || idle == 0ull) continue;
                const uint32_t avail = 256u - nxt_line[s];
                const uint32_t k = (uint32_t)__popcll(idle) < avail ? (uint32_t)__popcll(idle) : avail;
                const uint32_t rank = (uint32_t)__popcll(idle & lt_mask);
                if (my_slot < 0 && rank < k) {
                    my_slot = s;
                    const int line = (int)(nxt_line[s] + rank);
                    axis_geom((int)ax[s], line, base, stride);
                    lch = job.llr + (size_t)cube[s] * kCube;
                    le = job.le_scratch + (wave_id * S + s) * kCube;
                    soft = job.soft ? job.soft + (size_t)cube[s] * kCube : nullptr;
                    first = (it[s] == 1 && ax[s] == 0);
                    fresh = true;
                }
                nxt_line[s] += k;
                idle = __ballot(my_slot < 0);
            }
        }
        // ---- 4. set-up of newly assigned lines ----
        if (fresh) {
            const PSrc src{lch, le, job.alpha, base, stride, first};
            if (line_setup(src, L, hd, s0, pm_hd, ic)) {
                search_init(ss, ls, hd, s0, pm_hd, ic, cfg);
            } else {
                atomicOr(&cnt[4 * my_slot + 3], 1u);
                ss.done = true;
                ss.T = 0;
                ls.nl = 0;
                ls.pnl = 0.0;
            }
        }
        // ---- 5. a bounded step of every busy lane's guessing loop ----
        if (my_slot >= 0 && !ss.done)
            search_run<FINITE_T>(ss, ls, L, hd, s0, pm_hd, ic, cfg, step_len);
        // ---- 6. finished lines: soft output, L_E, hard decisions, counters ----
        if (my_slot >= 0 && ss.done) {
            const PSrc src{lch, le, job.alpha, base, stride, first};
            const double pnotin = fmax(ls.pnl, 1e-40);
            uint32_t *hb = hd_bits + 128 * my_slot;
#pragma unroll 1
            for (int i = 0; i < 16; ++i) {
                const int p = base + i * stride;
                const double xin = src(i);
                double la, lenew;
                if (ls.nl == 0) { // empty list: L_APP = llr, L_E = 0 (:537-539)
                    la = xin;
                    lenew = 0.0;
                } else {
                    la = soft_app(ls, pnotin, i, xin);
                    lenew = la - xin;
                }
                le[p] = lenew;
                if (soft) soft[p] = la;
                if (!(la > 0.0)) atomicOr(&hb[p >> 5], 1u << (p & 31));
            }
            atomicAdd(&cnt[4 * my_slot + 1], ss.T);
            atomicMax(&cnt[4 * my_slot + 2], ss.T);
            atomicAdd(&cnt[4 * my_slot + 0], 1u);
            my_slot = -1;
        }
    }
}

hipError_t launch_persistent(const CubeJob &job, int waves_per_block, int blocks, int slots,
                             hipStream_t st)
{
    if (job.n_cw == 0) return hipSuccess;
    if (waves_per_block != kWaves) return hipErrorInvalidValue;
    const bool finite = job.tmax < (1ull << 17);
    const uint32_t thresh = 16, step = 64;
    const dim3 g(blocks), b(64 * kWaves);
#define SG_LAUNCH(FT, SS)                                                                     \
    hipLaunchKernelGGL((persistent_kernel<FT, SS>), g, b, 0, st, job, thresh, step)
    if (slots == 1) {
        if (finite) SG_LAUNCH(true, 1); else SG_LAUNCH(false, 1);
    } else if (slots == 2) {
        if (finite) SG_LAUNCH(true, 2); else SG_LAUNCH(false, 2);
    } else if (slots == 3) {
        if (finite) SG_LAUNCH(true, 3); else SG_LAUNCH(false, 3);
    } else if (slots == 4) {
        if (finite) SG_LAUNCH(true, 4); else SG_LAUNCH(false, 4);
    } else {
        return hipErrorInvalidValue;
    }
#undef SG_LAUNCH
    return hipGetLastError();
}

} // namespace sg
