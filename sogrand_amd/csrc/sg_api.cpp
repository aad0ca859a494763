// C ABI of the SOGRAND gfx950 decoder (include/sogrand.h).
#include "sogrand.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "sg_common.h"
#include "sg_ctx.h"
#include "sg_launch.h"
#include "sg_schedule.h"

#pragma clang fp contract(off)

namespace sg {

bool is_device_ptr(const void *ptr)
{
    if (!ptr) return false;
    hipPointerAttribute_t a;
    hipError_t e = hipPointerGetAttributes(&a, ptr);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

} // namespace sg

using sg::effective_tmax;
using sg::is_device_ptr;

namespace sg {

int arith_of(const sogrand_params &p)
{
    if (p.variant == SOGRAND_VARIANT_LEGACY) return kArithF64Legacy;
    return p.precision == SOGRAND_FP32_MIXED ? kArithF32 : kArithF64;
}

// EPSILON of the confidence test and clamps, and the floor inside the soft
// output's logs: `epsilon` for both (cubic_decoder_final.c:16, :573), or
// cubic_decoder.c's 1e-9 (:503, :509-510, :710) and 1e-30 (:533).
static void eps_of(const sogrand_params &p, double &eps, double &eps_log)
{
    eps = p.variant == SOGRAND_VARIANT_LEGACY ? 1e-9 : p.epsilon;
    eps_log = p.variant == SOGRAND_VARIANT_LEGACY ? 1e-30 : p.epsilon;
}

LaunchPlan plan_launch(const sogrand_ctx *ctx, const sogrand_params &p, size_t n)
{
    LaunchPlan lp{};
    lp.persistent = p.backend == SOGRAND_BACKEND_PERSISTENT;
    lp.lanegroup = p.backend == SOGRAND_BACKEND_LANEGROUP;
    int auto_slots = 2;
    if (p.backend == SOGRAND_BACKEND_AUTO) {
        // W = the persistent waves resident at two cube slots each (3,072 on
        // MI355X). Up to W cubes a persistent wave would hold one cube with 64
        // lanes where the plain kernel gives it 256, and its queue has nothing
        // to balance: plain (6.7 against 27.3 ms for 256 codewords at 2 dB).
        // Up to 2W, one slot per wave fills the device sooner than two
        // (4,096 codewords: 37.3 against 38.9 ms at 2 dB, 171 against 229 at
        // 1 dB); from there on two slots (16,384: 94.9 against 104.2 ms).
        const uint64_t tmax = effective_tmax(p.Tmax);
        const size_t w = (size_t)ctx->cu_count * kPersistentWaves *
                         persistent_blocks_per_cu(2, tmax_binds(tmax), arith_of(p));
        lp.persistent = n > w;
        auto_slots = n <= 2 * w ? 1 : 2;
    }
    if (lp.lanegroup) {
        // the cube lives in LDS: no L_E scratch
        lp.blocks = (int)lanegroup_blocks(n, ctx->cu_count);
        lp.slots = 1;
        lp.le_slots = 0;
    } else if (lp.persistent) {
        const int slots = p.wave_slots > 0 ? p.wave_slots : auto_slots;
        const int wpb = kPersistentWaves;
        // a persistent grid must be resident all at once: never more workgroups
        // per CU than the occupancy the kernel's registers and LDS allow
        const uint64_t tmax = effective_tmax(p.Tmax);
        const int occ = persistent_blocks_per_cu(slots, tmax_binds(tmax), arith_of(p));
        int bpc = p.waves_per_cu > 0 ? (p.waves_per_cu + wpb - 1) / wpb : occ;
        if (bpc > occ) bpc = occ;
        int blocks = ctx->cu_count * bpc;
        // never more workers than cubes
        const size_t max_blocks = (n + (size_t)wpb * slots - 1) / ((size_t)wpb * slots) + 1;
        if ((size_t)blocks > max_blocks) blocks = (int)max_blocks;
        lp.blocks = blocks;
        lp.slots = slots;
        lp.le_slots = (size_t)blocks * wpb * slots * kPersistentSlotScratch;
    } else {
        lp.blocks = (int)plain_scratch_slots(n);
        lp.slots = 1;
        lp.le_slots = (size_t)lp.blocks;
    }
    return lp;
}

size_t max_le_slots(size_t n)
{
    // persistent: at most ceil(n / (waves x slots)) + 1 workgroups, i.e. at most
    // n + 2 x waves x slots cube slots (slots <= 4), each with its scratch cubes
    return kPersistentSlotScratch * (n + 2 * (size_t)kPersistentWaves * 4);
}

void fill_job(CubeJob &job, const sogrand_ctx *ctx, const sogrand_params &p)
{
    job.sched = ctx->d_sched;
    job.tmax = effective_tmax(p.Tmax);
    job.thres = 1.0 - p.p_ET;
    job.alpha = p.alpha;
    job.L = p.L;
    job.Imax = p.Imax;
    job.arith = arith_of(p);
    eps_of(p, job.eps, job.eps_log);
}

hipError_t launch_cubes(const sogrand_params &p, const LaunchPlan &lp, const CubeJob &job,
                        hipStream_t st, hipEvent_t ev_start)
{
    if (lp.persistent) {
        hipError_t e = launch_fill_u32(job.queue, 0u, 1, st);
        if (e != hipSuccess) return e;
        if (ev_start) (void)hipEventRecord(ev_start, st);
        const uint32_t batch = p.lane_batch > 0 ? (uint32_t)p.lane_batch : 4u;
        const uint32_t step = p.step_len > 0 ? (uint32_t)p.step_len : 64u;
        return launch_persistent(job, kPersistentWaves, lp.blocks, lp.slots, batch, step, st);
    }
    if (ev_start) (void)hipEventRecord(ev_start, st);
    if (lp.lanegroup) return launch_lanegroup(job, lp.blocks, st);
    return launch_plain(job, lp.blocks, st);
}

} // namespace sg

extern "C" {

void sogrand_params_default(sogrand_params *p)
{
    if (!p) return;
    std::memset(p, 0, sizeof *p);
    p->L = 3;
    p->Imax = 30;
    p->Tmax = UINT64_MAX;
    p->p_ET = 1e-5;
    p->alpha = 0.7;
    p->precision = SOGRAND_FP64;
    p->backend = SOGRAND_BACKEND_PERSISTENT;
    p->wave_slots = 0;
    p->waves_per_cu = 0;
    p->lane_batch = 0;
    p->step_len = 0;
    p->use_graph = 0;
    p->epsilon = 1e-40;
    p->variant = SOGRAND_VARIANT_FINAL;
}

const char *sogrand_strerror(int code)
{
    switch (code) {
    case SOGRAND_OK: return "ok";
    case SOGRAND_EINVAL: return "invalid argument";
    case SOGRAND_ENOMEM: return "out of memory";
    case SOGRAND_EHIP: return "HIP runtime error";
    case SOGRAND_ENONFINITE: return "non-finite LLR (no longer returned)";
    case SOGRAND_ENODEV: return "no gfx950 HIP device";
    case SOGRAND_EIO: return "file I/O error";
    case SOGRAND_EBUSY: return "admission refused";
    default: return "unknown error";
    }
}

static int check_params(const sogrand_params *p)
{
    if (!p) return SOGRAND_EINVAL;
    if (p->L < 1 || p->L > SOGRAND_MAX_L) return SOGRAND_EINVAL;
    if (p->Imax < 1 || p->Imax > 1000) return SOGRAND_EINVAL;
    if (!(p->p_ET >= 0.0 && p->p_ET <= 1.0)) return SOGRAND_EINVAL;
    if (!std::isfinite(p->alpha)) return SOGRAND_EINVAL;
    if (p->precision != SOGRAND_FP64 && p->precision != SOGRAND_FP32_MIXED) return SOGRAND_EINVAL;
    if (p->backend != SOGRAND_BACKEND_PLAIN && p->backend != SOGRAND_BACKEND_PERSISTENT &&
        p->backend != SOGRAND_BACKEND_LANEGROUP && p->backend != SOGRAND_BACKEND_AUTO)
        return SOGRAND_EINVAL;
    // the lane-group kernel has the final decoder's double-precision arithmetic only
    if (p->backend == SOGRAND_BACKEND_LANEGROUP &&
        (p->precision != SOGRAND_FP64 || p->variant != SOGRAND_VARIANT_FINAL))
        return SOGRAND_EINVAL;
    if (p->wave_slots < 0 || p->wave_slots > 4 || p->waves_per_cu < 0 || p->waves_per_cu > 32)
        return SOGRAND_EINVAL;
    if (p->lane_batch < 0 || p->lane_batch > 64 || p->step_len < 0 || p->step_len > 65536)
        return SOGRAND_EINVAL;
    if (!(p->epsilon > 0.0 && p->epsilon < 0.5)) return SOGRAND_EINVAL;
    if (p->variant != SOGRAND_VARIANT_FINAL && p->variant != SOGRAND_VARIANT_LEGACY)
        return SOGRAND_EINVAL;
    if (p->variant == SOGRAND_VARIANT_LEGACY && p->precision != SOGRAND_FP64) return SOGRAND_EINVAL;
    return SOGRAND_OK;
}

int sogrand_ctx_create(sogrand_ctx **out, int device, const sogrand_params *params)
{
    if (!out) return SOGRAND_EINVAL;
    *out = nullptr;
    sogrand_params p;
    if (params)
        p = *params;
    else
        sogrand_params_default(&p);
    int rc = check_params(&p);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
        (void)hipGetLastError();
        return SOGRAND_ENODEV;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return SOGRAND_EHIP;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        std::fprintf(stderr, "sogrand: device %d is %s; this library is built for gfx950 only\n",
                     device, prop.gcnArchName);
        return SOGRAND_ENODEV;
    }
    sogrand_ctx *c = new (std::nothrow) sogrand_ctx;
    if (!c) return SOGRAND_ENOMEM;
    c->device = device;
    c->p = p;
    c->cu_count = prop.multiProcessorCount;
    if (hipSetDevice(device) != hipSuccess) {
        delete c;
        return SOGRAND_EHIP;
    }
    if (hipStreamCreate(&c->own_stream) != hipSuccess || hipEventCreate(&c->ev0) != hipSuccess ||
        hipEventCreate(&c->ev1) != hipSuccess) {
        sogrand_ctx_destroy(c);
        return SOGRAND_EHIP;
    }
    // mask tables and position-word tables in one buffer (sg_common.h)
    std::vector<uint8_t> buf(sg::kSchedBufferBytes, 0);
    uint16_t *masks = reinterpret_cast<uint16_t *>(buf.data());
    uint32_t *words = reinterpret_cast<uint32_t *>(buf.data() + sg::kSchedPosOffsetBytes);
    sg::build_all_schedule_tables(masks);
    for (int t = 0; t < sg::kSchedTables; ++t) {
        const size_t o = (size_t)t * sg::kSchedStride;
        for (size_t j = 0; j < (size_t)sg::kSchedStride; ++j) {
            const bool level = j == 0 || sg::pattern_wt(t, masks[o + j]) !=
                                             sg::pattern_wt(t, masks[o + j - 1]);
            words[o + j] = sg::pos_word(masks[o + j], level);
        }
    }
    if (hipMalloc(&c->d_sched, buf.size()) != hipSuccess) {
        sogrand_ctx_destroy(c);
        return SOGRAND_ENOMEM;
    }
    if (hipMemcpy(c->d_sched, buf.data(), buf.size(), hipMemcpyHostToDevice) != hipSuccess) {
        sogrand_ctx_destroy(c);
        return SOGRAND_EHIP;
    }
    // the no-allocation entry point's scratch on this device (sg_noalloc.hip)
    if (sg::noalloc_scratch_init(device) != hipSuccess) {
        sogrand_ctx_destroy(c);
        return SOGRAND_EHIP;
    }
    *out = c;
    return SOGRAND_OK;
}

int sogrand_ctx_set_params(sogrand_ctx *ctx, const sogrand_params *params)
{
    if (!ctx) return SOGRAND_EINVAL;
    int rc = check_params(params);
    if (rc) return rc;
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->p = *params;
    return SOGRAND_OK;
}

void sogrand_ctx_destroy(sogrand_ctx *ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->own_stream) (void)hipStreamSynchronize(ctx->own_stream);
    ctx->le_scratch.release();
    ctx->llr_stage.release();
    ctx->soft_stage.release();
    ctx->sigma_stage.release();
    ctx->info_stage.release();
    ctx->stats_stage.release();
    ctx->queue.release();
    ctx->nguess_stage.release();
    for (HostStage &h : ctx->hs) {
        if (h.st) (void)hipStreamSynchronize(h.st);
        h.llr.release();
        h.soft.release();
        h.le.release();
        h.info.release();
        h.stats.release();
        h.queue.release();
        if (h.ev0) (void)hipEventDestroy(h.ev0);
        if (h.ev1) (void)hipEventDestroy(h.ev1);
        if (h.st) (void)hipStreamDestroy(h.st);
    }
    if (ctx->d_sched) (void)hipFree(ctx->d_sched);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->graph) (void)hipGraphExecDestroy(ctx->graph);
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
    delete ctx;
}


// Decode one chunk whose buffers are all on the device.
static int decode_device_chunk(sogrand_ctx *ctx, const double *d_llr, size_t n, uint8_t *d_info,
                               double *d_soft, uint32_t *d_stats, hipStream_t st, float *ms,
                               bool use_graph = false)
{
    const sogrand_params &p = ctx->p;
    sg::CubeJob job{};
    job.llr = d_llr;
    job.info = d_info;
    job.soft = d_soft;
    job.stats = d_stats;
    job.n_cw = n;
    sg::fill_job(job, ctx, p);
    const sg::LaunchPlan lp = sg::plan_launch(ctx, p, n);
    if (ctx->le_scratch.reserve(lp.le_slots * sg::kCube) != hipSuccess) return SOGRAND_ENOMEM;
    if (ctx->queue.reserve(4) != hipSuccess) return SOGRAND_ENOMEM;
    job.le_scratch = ctx->le_scratch.p;
    job.queue = ctx->queue.p;
    if (use_graph) {
        // the captured sequence is the queue-reset fill kernel and the decode
        // kernel (no memset nodes); recapture whenever buffers, size, stream or
        // parameters change
        GraphKey k;
        std::memset(&k, 0, sizeof k); // compared bytewise, padding included
        k.llr = d_llr;
        k.info = d_info;
        k.soft = d_soft;
        k.stats = d_stats;
        k.le = job.le_scratch;
        k.queue = job.queue;
        k.n = n;
        k.st = st;
        std::memcpy(&k.p, &p, sizeof p);
        if (!ctx->graph || std::memcmp(&k, &ctx->graph_key, sizeof k) != 0) {
            if (ctx->graph) (void)hipGraphExecDestroy(ctx->graph);
            ctx->graph = nullptr;
            hipGraph_t g = nullptr;
            if (hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed) != hipSuccess)
                return SOGRAND_EHIP;
            const hipError_t e1 = sg::launch_cubes(p, lp, job, st, nullptr);
            const hipError_t e2 = hipStreamEndCapture(st, &g);
            if (e1 != hipSuccess || e2 != hipSuccess || !g) {
                if (g) (void)hipGraphDestroy(g);
                return SOGRAND_EHIP;
            }
            const hipError_t e3 = hipGraphInstantiate(&ctx->graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e3 != hipSuccess) {
                ctx->graph = nullptr;
                return SOGRAND_EHIP;
            }
            std::memcpy(&ctx->graph_key, &k, sizeof k);
            ctx->graph_replays = 0;
        } else {
            ctx->graph_replays += 1;
        }
        (void)hipEventRecord(ctx->ev0, st);
        if (hipGraphLaunch(ctx->graph, st) != hipSuccess) return SOGRAND_EHIP;
    } else if (sg::launch_cubes(p, lp, job, st, ctx->ev0) != hipSuccess) {
        return SOGRAND_EHIP;
    }
    (void)hipEventRecord(ctx->ev1, st);
    if (hipEventSynchronize(ctx->ev1) != hipSuccess) return SOGRAND_EHIP;
    float t = 0.f;
    (void)hipEventElapsedTime(&t, ctx->ev0, ctx->ev1);
    *ms += t;
    return SOGRAND_OK;
}

// Results of the chunk in flight on stage h to the caller's host buffers (or,
// for device-resident outputs, nothing to copy), after its stream has finished.
static int finish_stage(sogrand_ctx *ctx, HostStage &h, uint8_t *info_bits, double *soft_out,
                        sogrand_cw_stats *stats, bool dev_info, bool dev_soft, bool dev_stats)
{
    if (!h.busy) return SOGRAND_OK;
    h.busy = false;
    if (hipStreamSynchronize(h.st) != hipSuccess) return SOGRAND_EHIP;
    const size_t off = h.off, n = h.n;
    if (!dev_info && hipMemcpy(info_bits + off * sg::kInfoBytes, h.info.p, n * sg::kInfoBytes,
                               hipMemcpyDeviceToHost) != hipSuccess)
        return SOGRAND_EHIP;
    if (soft_out && !dev_soft &&
        hipMemcpy(soft_out + off * sg::kCube, h.soft.p, n * sg::kCube * sizeof(double),
                  hipMemcpyDeviceToHost) != hipSuccess)
        return SOGRAND_EHIP;
    if (stats && !dev_stats &&
        hipMemcpy(stats + off, h.stats.p, n * sizeof(sogrand_cw_stats), hipMemcpyDeviceToHost) !=
            hipSuccess)
        return SOGRAND_EHIP;
    (void)ctx;
    return SOGRAND_OK;
}

// Decode with some buffers in host memory: balanced chunks of at most 32,768
// codewords alternate between two stages, so that each chunk's input copy and
// launch overlap the previous chunk's kernel (and its tail), and each stage's
// results are copied out when the stage is next needed.
static int decode_host_pipelined(sogrand_ctx *ctx, const double *llr, size_t n_cw,
                                 uint8_t *info_bits, double *soft_out, sogrand_cw_stats *stats,
                                 bool dev_llr, bool dev_info, bool dev_soft, bool dev_stats,
                                 float *ms)
{
    const sogrand_params &p = ctx->p;
    const size_t max_chunk = 32768;
    const size_t n_chunks = (n_cw + max_chunk - 1) / max_chunk;
    const size_t chunk = (n_cw + n_chunks - 1) / n_chunks;
    for (HostStage &h : ctx->hs) {
        if (!h.st && hipStreamCreateWithFlags(&h.st, hipStreamNonBlocking) != hipSuccess) {
            h.st = nullptr;
            return SOGRAND_EHIP;
        }
        if (!h.ev0 && hipEventCreate(&h.ev0) != hipSuccess) return SOGRAND_EHIP;
        if (!h.ev1 && hipEventCreate(&h.ev1) != hipSuccess) return SOGRAND_EHIP;
        h.busy = false; // every call finishes both stages before it returns
    }
    int rc = SOGRAND_OK;
    size_t k = 0;
    for (size_t off = 0; off < n_cw; off += chunk, ++k) {
        HostStage &h = ctx->hs[k & 1];
        rc = finish_stage(ctx, h, info_bits, soft_out, stats, dev_info, dev_soft, dev_stats);
        if (rc != SOGRAND_OK) break;
        const size_t n = (n_cw - off) < chunk ? (n_cw - off) : chunk;
        const double *d_llr = llr + off * sg::kCube;
        if (!dev_llr) {
            if (h.llr.reserve(n * sg::kCube) != hipSuccess) { rc = SOGRAND_ENOMEM; break; }
            if (hipMemcpyAsync(h.llr.p, d_llr, n * sg::kCube * sizeof(double),
                               hipMemcpyHostToDevice, h.st) != hipSuccess)
                { rc = SOGRAND_EHIP; break; }
            d_llr = h.llr.p;
        }
        uint8_t *d_info = info_bits + off * sg::kInfoBytes;
        if (!dev_info) {
            if (h.info.reserve(n * sg::kInfoBytes) != hipSuccess) { rc = SOGRAND_ENOMEM; break; }
            d_info = h.info.p;
        }
        double *d_soft = nullptr;
        if (soft_out) {
            d_soft = soft_out + off * sg::kCube;
            if (!dev_soft) {
                if (h.soft.reserve(n * sg::kCube) != hipSuccess) { rc = SOGRAND_ENOMEM; break; }
                d_soft = h.soft.p;
            }
        }
        uint32_t *d_stats = nullptr;
        if (stats) {
            d_stats = (uint32_t *)(stats + off);
            if (!dev_stats) {
                if (h.stats.reserve(n * 6) != hipSuccess) { rc = SOGRAND_ENOMEM; break; }
                d_stats = h.stats.p;
            }
        }
        sg::CubeJob job{};
        job.llr = d_llr;
        job.info = d_info;
        job.soft = d_soft;
        job.stats = d_stats;
        job.n_cw = n;
        sg::fill_job(job, ctx, p);
        const sg::LaunchPlan lp = sg::plan_launch(ctx, p, n);
        if (h.le.reserve(lp.le_slots * sg::kCube) != hipSuccess) { rc = SOGRAND_ENOMEM; break; }
        if (h.queue.reserve(4) != hipSuccess) { rc = SOGRAND_ENOMEM; break; }
        job.le_scratch = h.le.p;
        job.queue = h.queue.p;
        if (sg::launch_cubes(p, lp, job, h.st, k == 0 ? ctx->ev0 : h.ev0) != hipSuccess) {
            rc = SOGRAND_EHIP;
            break;
        }
        if (hipEventRecord(h.ev1, h.st) != hipSuccess) { rc = SOGRAND_EHIP; break; }
        h.off = off;
        h.n = n;
        h.busy = true;
        ctx->last_launches += 1;
    }
    // finish both stages whatever happened, the later chunk last
    for (size_t j = 0; j < 2; ++j) {
        HostStage &h = ctx->hs[(k + j) & 1];
        const int r = finish_stage(ctx, h, info_bits, soft_out, stats, dev_info, dev_soft,
                                   dev_stats);
        if (rc == SOGRAND_OK) rc = r;
    }
    // an error path can leave a copy enqueued on a stage that never launched
    for (HostStage &h : ctx->hs) (void)hipStreamSynchronize(h.st);
    // kernel time: from the first kernel's start to the last kernel's end (the
    // launches overlap, so their durations do not add up to it)
    if (rc == SOGRAND_OK && k > 0) {
        float span = 0.f;
        for (size_t j = 0; j < 2 && j < k; ++j) {
            float t = 0.f;
            if (hipEventElapsedTime(&t, ctx->ev0, ctx->hs[j].ev1) == hipSuccess && t > span)
                span = t;
        }
        *ms += span;
    }
    return rc;
}

int sogrand_decode_cubic(sogrand_ctx *ctx, const double *llr, size_t n_cw, uint8_t *info_bits,
                         double *soft_out, sogrand_cw_stats *stats, void *hip_stream)
{
    if (!ctx || (!llr && n_cw) || (!info_bits && n_cw)) return SOGRAND_EINVAL;
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->last_ms = 0.0;
    ctx->last_launches = 0;
    if (n_cw == 0) return SOGRAND_OK;
    if (hipSetDevice(ctx->device) != hipSuccess) return SOGRAND_EHIP;
    hipStream_t st = hip_stream ? (hipStream_t)hip_stream : ctx->own_stream;
    static_assert(sizeof(sogrand_cw_stats) == 24, "stats layout");

    const bool dev_llr = is_device_ptr(llr), dev_info = is_device_ptr(info_bits);
    const bool dev_soft = soft_out && is_device_ptr(soft_out);
    const bool dev_stats = stats && is_device_ptr(stats);
    const bool all_dev = dev_llr && dev_info && (!soft_out || dev_soft) && (!stats || dev_stats);
    float ms = 0.f;
    int rc = SOGRAND_OK;
    if (all_dev) {
        rc = decode_device_chunk(ctx, llr, n_cw, info_bits, soft_out, (uint32_t *)stats, st, &ms,
                                 ctx->p.use_graph != 0);
        ctx->last_launches = 1;
    } else if (!hip_stream) {
        rc = decode_host_pipelined(ctx, llr, n_cw, info_bits, soft_out, stats, dev_llr, dev_info,
                                   dev_soft, dev_stats, &ms);
    } else {
        // on the caller's stream: one chunk at a time, in order
        const size_t chunk = 65536;
        for (size_t off = 0; off < n_cw && rc == SOGRAND_OK; off += chunk) {
            const size_t n = (n_cw - off) < chunk ? (n_cw - off) : chunk;
            const double *d_llr = llr + off * sg::kCube;
            if (!dev_llr) {
                if (ctx->llr_stage.reserve(n * sg::kCube) != hipSuccess) return SOGRAND_ENOMEM;
                if (hipMemcpyAsync(ctx->llr_stage.p, d_llr, n * sg::kCube * sizeof(double),
                                   hipMemcpyHostToDevice, st) != hipSuccess)
                    return SOGRAND_EHIP;
                d_llr = ctx->llr_stage.p;
            }
            uint8_t *d_info = info_bits + off * sg::kInfoBytes;
            if (!dev_info) {
                if (ctx->info_stage.reserve(n * sg::kInfoBytes) != hipSuccess) return SOGRAND_ENOMEM;
                d_info = ctx->info_stage.p;
            }
            double *d_soft = nullptr;
            if (soft_out) {
                d_soft = soft_out + off * sg::kCube;
                if (!dev_soft) {
                    if (ctx->soft_stage.reserve(n * sg::kCube) != hipSuccess) return SOGRAND_ENOMEM;
                    d_soft = ctx->soft_stage.p;
                }
            }
            uint32_t *d_stats = nullptr;
            if (stats) {
                d_stats = (uint32_t *)(stats + off);
                if (!dev_stats) {
                    if (ctx->stats_stage.reserve(n * 6) != hipSuccess) return SOGRAND_ENOMEM;
                    d_stats = ctx->stats_stage.p;
                }
            }
            rc = decode_device_chunk(ctx, d_llr, n, d_info, d_soft, d_stats, st, &ms);
            ctx->last_launches += 1;
            if (rc != SOGRAND_OK) break;
            if (!dev_info &&
                hipMemcpy(info_bits + off * sg::kInfoBytes, d_info, n * sg::kInfoBytes,
                          hipMemcpyDeviceToHost) != hipSuccess)
                return SOGRAND_EHIP;
            if (soft_out && !dev_soft &&
                hipMemcpy(soft_out + off * sg::kCube, d_soft, n * sg::kCube * sizeof(double),
                          hipMemcpyDeviceToHost) != hipSuccess)
                return SOGRAND_EHIP;
            if (stats && !dev_stats &&
                hipMemcpy(stats + off, d_stats, n * sizeof(sogrand_cw_stats),
                          hipMemcpyDeviceToHost) != hipSuccess)
                return SOGRAND_EHIP;
        }
    }
    ctx->last_ms = ms;
    return rc;
}

int sogrand_decode_square(sogrand_ctx *ctx, const double *llr, size_t n_cw, uint8_t *info_bits,
                          double *soft_out, sogrand_cw_stats *stats, void *hip_stream)
{
    if (!ctx || (!llr && n_cw) || (!info_bits && n_cw)) return SOGRAND_EINVAL;
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->last_ms = 0.0;
    ctx->last_launches = 0;
    if (n_cw == 0) return SOGRAND_OK;
    if (hipSetDevice(ctx->device) != hipSuccess) return SOGRAND_EHIP;
    hipStream_t st = hip_stream ? (hipStream_t)hip_stream : ctx->own_stream;
    constexpr size_t kPos = 256, kInfo = 8;
    // host buffers are staged through the context's device buffers whole (a
    // square is 2 KB of LLRs: a million squares stage 2 GB)
    const double *d_llr = llr;
    if (!is_device_ptr(llr)) {
        if (ctx->llr_stage.reserve(n_cw * kPos) != hipSuccess) return SOGRAND_ENOMEM;
        if (hipMemcpyAsync(ctx->llr_stage.p, llr, n_cw * kPos * sizeof(double),
                           hipMemcpyHostToDevice, st) != hipSuccess)
            return SOGRAND_EHIP;
        d_llr = ctx->llr_stage.p;
    }
    const bool dev_info = is_device_ptr(info_bits);
    uint8_t *d_info = info_bits;
    if (!dev_info) {
        if (ctx->info_stage.reserve(n_cw * kInfo) != hipSuccess) return SOGRAND_ENOMEM;
        d_info = ctx->info_stage.p;
    }
    const bool dev_soft = soft_out && is_device_ptr(soft_out);
    double *d_soft = soft_out;
    if (soft_out && !dev_soft) {
        if (ctx->soft_stage.reserve(n_cw * kPos) != hipSuccess) return SOGRAND_ENOMEM;
        d_soft = ctx->soft_stage.p;
    }
    const bool dev_stats = stats && is_device_ptr(stats);
    uint32_t *d_stats = (uint32_t *)stats;
    if (stats && !dev_stats) {
        if (ctx->stats_stage.reserve(n_cw * 6) != hipSuccess) return SOGRAND_ENOMEM;
        d_stats = ctx->stats_stage.p;
    }
    sg::CubeJob job{};
    job.llr = d_llr;
    job.info = d_info;
    job.soft = d_soft;
    job.stats = d_stats;
    job.n_cw = n_cw;
    sg::fill_job(job, ctx, ctx->p);
    const size_t blocks = sg::square_blocks(n_cw);
    if (ctx->le_scratch.reserve(blocks * sg::kCube) != hipSuccess) return SOGRAND_ENOMEM;
    job.le_scratch = ctx->le_scratch.p;
    (void)hipEventRecord(ctx->ev0, st);
    if (sg::launch_square(job, (int)blocks, st) != hipSuccess) return SOGRAND_EHIP;
    (void)hipEventRecord(ctx->ev1, st);
    if (hipEventSynchronize(ctx->ev1) != hipSuccess) return SOGRAND_EHIP;
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    ctx->last_ms = ms;
    ctx->last_launches = 1;
    if (!dev_info && hipMemcpy(info_bits, d_info, n_cw * kInfo, hipMemcpyDeviceToHost) != hipSuccess)
        return SOGRAND_EHIP;
    if (soft_out && !dev_soft &&
        hipMemcpy(soft_out, d_soft, n_cw * kPos * sizeof(double), hipMemcpyDeviceToHost) !=
            hipSuccess)
        return SOGRAND_EHIP;
    if (stats && !dev_stats &&
        hipMemcpy(stats, d_stats, n_cw * sizeof(sogrand_cw_stats), hipMemcpyDeviceToHost) !=
            hipSuccess)
        return SOGRAND_EHIP;
    return SOGRAND_OK;
}

int sogrand_bitso_lines(sogrand_ctx *ctx, const double *llr, size_t n_lines, double *L_APP,
                        double *L_E, uint32_t *N_guess, void *hip_stream)
{
    if (!ctx || (!llr && n_lines)) return SOGRAND_EINVAL;
    std::lock_guard<std::mutex> g(ctx->mu);
    if (n_lines == 0) return SOGRAND_OK;
    if (hipSetDevice(ctx->device) != hipSuccess) return SOGRAND_EHIP;
    hipStream_t st = hip_stream ? (hipStream_t)hip_stream : ctx->own_stream;
    // stage everything through one device buffer: llr | la | le
    if (ctx->llr_stage.reserve(n_lines * 16 * 3) != hipSuccess) return SOGRAND_ENOMEM;
    if (ctx->nguess_stage.reserve(n_lines) != hipSuccess) return SOGRAND_ENOMEM;
    double *d_in = ctx->llr_stage.p, *d_la = d_in + n_lines * 16, *d_le = d_la + n_lines * 16;
    if (hipMemcpyAsync(d_in, llr, n_lines * 16 * sizeof(double), hipMemcpyDefault, st) !=
        hipSuccess)
        return SOGRAND_EHIP;
    double eps, eps_log;
    sg::eps_of(ctx->p, eps, eps_log);
    if (sg::launch_lines(d_in, n_lines, d_la, d_le, ctx->nguess_stage.p, ctx->d_sched, effective_tmax(ctx->p.Tmax), 1.0 - ctx->p.p_ET, ctx->p.L,
                         eps, eps_log, sg::arith_of(ctx->p), st) != hipSuccess)
        return SOGRAND_EHIP;
    if (hipStreamSynchronize(st) != hipSuccess) return SOGRAND_EHIP;
    if (L_APP && hipMemcpy(L_APP, d_la, n_lines * 16 * sizeof(double), hipMemcpyDefault) !=
                     hipSuccess)
        return SOGRAND_EHIP;
    if (L_E &&
        hipMemcpy(L_E, d_le, n_lines * 16 * sizeof(double), hipMemcpyDefault) != hipSuccess)
        return SOGRAND_EHIP;
    if (N_guess && hipMemcpy(N_guess, ctx->nguess_stage.p, n_lines * sizeof(uint32_t),
                             hipMemcpyDefault) != hipSuccess)
        return SOGRAND_EHIP;
    return SOGRAND_OK;
}

int sogrand_graph_replays(sogrand_ctx *ctx, uint32_t *replays)
{
    if (!ctx || !replays) return SOGRAND_EINVAL;
    std::lock_guard<std::mutex> g(ctx->mu);
    *replays = ctx->graph ? ctx->graph_replays : 0u;
    return SOGRAND_OK;
}

int sogrand_last_kernel_ms(sogrand_ctx *ctx, double *ms, uint32_t *launches)
{
    if (!ctx) return SOGRAND_EINVAL;
    if (ms) *ms = ctx->last_ms;
    if (launches) *launches = ctx->last_launches;
    return SOGRAND_OK;
}

int sogrand_schedule_table(int32_t ic, uint16_t *out)
{
    if (!out || ic < 0) return SOGRAND_EINVAL;
    sg::build_schedule_table(ic < sg::kSchedGenericIc ? ic : sg::kSchedGenericIc, out);
    return SOGRAND_OK;
}

// ---- synthetic data ---------------------------------------------------------

double sogrand_sigma_for_ebn0(double ebn0_db)
{
    const double R = std::pow(8.0 / 16.0, 3.0); // channel_sim.c:32
    const double esn0 = ebn0_db + 10.0 * std::log10(2.0 * R);
    return 1.0 / std::sqrt(std::pow(10.0, esn0 / 10.0));
}

int sogrand_gen_info_bits(uint64_t seed, uint64_t cw_offset, size_t n_cw, uint8_t *info_bits)
{
    if (!info_bits && n_cw) return SOGRAND_EINVAL;
    for (size_t c = 0; c < n_cw; ++c)
        for (int j = 0; j < 8; ++j) {
            const uint64_t w = sg::info_word(seed, cw_offset + c, j);
            for (int q = 0; q < 8; ++q) info_bits[c * 64 + 8 * j + q] = (uint8_t)(w >> (8 * q));
        }
    return SOGRAND_OK;
}

int sogrand_encode_cubes(const uint8_t *info_bits, size_t n_cw, uint8_t *code_bytes)
{
    if ((!info_bits || !code_bytes) && n_cw) return SOGRAND_EINVAL;
    uint32_t W[256];
    for (size_t c = 0; c < n_cw; ++c) {
        sg::encode_cube_words(info_bits + c * 64, W);
        uint8_t *o = code_bytes + c * 512;
        for (int w = 0; w < 256; ++w) { // bit idx = 16*w + a0, MSB first
            o[2 * w] = (uint8_t)sg::rev8(W[w] & 0xffu);
            o[2 * w + 1] = (uint8_t)sg::rev8((W[w] >> 8) & 0xffu);
        }
    }
    return SOGRAND_OK;
}

int sogrand_gen_llr(uint64_t seed, uint64_t cw_offset, size_t n_cw, const double *sigma,
                    size_t n_sigma, double *llr)
{
    if ((!sigma || !llr) && n_cw) return SOGRAND_EINVAL;
    if (n_cw && n_sigma != 1 && n_sigma != n_cw) return SOGRAND_EINVAL;
    uint8_t info[64];
    uint32_t W[256];
    for (size_t c = 0; c < n_cw; ++c) {
        const uint64_t cw = cw_offset + c;
        for (int j = 0; j < 8; ++j) {
            const uint64_t w = sg::info_word(seed, cw, j);
            for (int q = 0; q < 8; ++q) info[8 * j + q] = (uint8_t)(w >> (8 * q));
        }
        sg::encode_cube_words(info, W);
        const double sgm = sigma[n_sigma == 1 ? 0 : c];
        const double var = sgm * sgm;
        double *out = llr + c * sg::kCube;
        for (int k = 0; k < 2048; ++k) {
            double z0, z1;
            sg::normal_pair(seed, cw, (uint64_t)k, &z0, &z1);
            const int p0 = 2 * k, p1 = 2 * k + 1;
            const double x0 = ((W[p0 >> 4] >> (p0 & 15)) & 1u) ? -1.0 : 1.0;
            const double x1 = ((W[p1 >> 4] >> (p1 & 15)) & 1u) ? -1.0 : 1.0;
            const double n0 = sgm * z0, n1 = sgm * z1;
            const double y0 = x0 + n0, y1 = x1 + n1;
            out[p0] = 2.0 * y0 / var;
            out[p1] = 2.0 * y1 / var;
        }
    }
    return SOGRAND_OK;
}

int sogrand_gen_llr_device(sogrand_ctx *ctx, uint64_t seed, uint64_t cw_offset, size_t n_cw,
                           const double *sigma, size_t n_sigma, double *llr, void *hip_stream)
{
    if (!ctx || ((!sigma || !llr) && n_cw)) return SOGRAND_EINVAL;
    if (n_cw && n_sigma != 1 && n_sigma != n_cw) return SOGRAND_EINVAL;
    if (n_cw == 0) return SOGRAND_OK;
    std::lock_guard<std::mutex> g(ctx->mu);
    if (hipSetDevice(ctx->device) != hipSuccess) return SOGRAND_EHIP;
    hipStream_t st = hip_stream ? (hipStream_t)hip_stream : ctx->own_stream;
    if (!is_device_ptr(llr)) return SOGRAND_EINVAL;
    const double *d_sigma = sigma;
    if (!is_device_ptr(sigma)) {
        if (ctx->sigma_stage.reserve(n_sigma) != hipSuccess) return SOGRAND_ENOMEM;
        if (hipMemcpyAsync(ctx->sigma_stage.p, sigma, n_sigma * sizeof(double),
                           hipMemcpyHostToDevice, st) != hipSuccess)
            return SOGRAND_EHIP;
        d_sigma = ctx->sigma_stage.p;
    }
    if (sg::launch_gen(seed, cw_offset, n_cw, d_sigma, n_sigma, llr, st) != hipSuccess)
        return SOGRAND_EHIP;
    if (hipStreamSynchronize(st) != hipSuccess) return SOGRAND_EHIP;
    return SOGRAND_OK;
}

} // extern "C"
