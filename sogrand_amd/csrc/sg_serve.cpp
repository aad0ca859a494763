// Serving loop: continuous batching of decode requests (frames of codewords).
//
// A worker thread owns `max_inflight` batch slots. Each slot has its own
// non-blocking HIP stream and a fixed carve of the device-memory budget (LLRs,
// L_E scratch, bits and statistics of up to `max_batch_cw` codewords), so
// admission never allocates. Requests wait in a FIFO; whenever a slot is free,
// every waiting request that fits is gathered into one batch (input copies on
// the slot's stream), decoded by one launch of the configured kernel, and the
// results are copied back when the slot's completion event fires. Batches on
// different slots run concurrently on the device: a request that arrives
// while others decode is admitted as soon as a slot frees up, and its
// workgroups take compute units as the earlier launches drain.
//
// Each codeword's result depends only on its LLRs (the decoder is batch
// independent), so a request decodes to the same bits, statistics and soft
// output alone or in any mix.
//
// The reference has no serving path; its nearest relative is the multi-stream
// CUDA decoder (SOGRAND_CUDA/cubic_decoder.cu:1145-1179), which writes batches
// out of order. Here results are returned by request id.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "sg_common.h"
#include "sg_ctx.h"
#include "sg_launch.h"
#include "sogrand.h"

namespace {

struct Request {
    uint64_t id;
    const double *llr;
    size_t n;
    uint8_t *info;
    double *soft;
    sogrand_cw_stats *stats;
    double t_submit;
    bool llr_dev, info_dev, soft_dev, stats_dev; // device (not host) memory
};

struct Slot {
    hipStream_t st = nullptr;
    hipEvent_t done = nullptr;
    double *llr = nullptr, *le = nullptr, *soft = nullptr;
    uint8_t *info = nullptr;
    uint32_t *stats = nullptr, *queue = nullptr;
    void *base = nullptr; // one device allocation per slot
    // pinned host staging: results (info | stats) always, soft output
    // when reserved, host-resident inputs on first use
    uint8_t *h_out = nullptr;
    double *h_soft = nullptr;
    double *h_llr = nullptr;
    bool busy = false;
    bool launch_ok = true; // every call of the last launch() succeeded
    double t_start = 0.0;
    std::vector<Request> reqs;
};

} // namespace

struct sogrand_server {
    sogrand_ctx *ctx = nullptr;
    sogrand_params p{};
    sogrand_server_cfg cfg{};
    size_t batch_cw = 0;
    size_t plain_below = 0; // in-flight codewords below which batches run the plain kernel
    size_t hold_above = 0;  // no new batch while this many codewords are in flight
    size_t lanegroup_max = 0; // batches up to this size run the lane-group kernel
    size_t small_batch = 0;   // batch size while at most `backlog` codewords wait
    size_t backlog = 0;
    size_t inflight_cw = 0; // codewords of the batches on the device (worker thread only)
    std::vector<Slot> slots;
    std::deque<Request> pending;
    size_t pending_cw = 0;
    std::deque<sogrand_completion> completed;
    uint64_t submitted = 0, finished = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    bool stop = false;
    bool fatal = false;
    std::thread worker;
    std::chrono::steady_clock::time_point t0;

    double now() const
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    void run();
    void take(Slot &s);
    void launch(Slot &s);
    void finish(Slot &s, bool event_ok);
};

// Move waiting requests into slot s (caller holds the lock): every one that
// fits, or, while the queue is short, at most small_batch codewords (at least
// one request).
void sogrand_server::take(Slot &s)
{
    s.reqs.clear();
    size_t n = 0;
    // (never more than a slot holds)
    const size_t limit =
        small_batch && pending_cw <= backlog ? std::min(small_batch, batch_cw) : batch_cw;
    while (!pending.empty() &&
           (s.reqs.empty() ? pending.front().n <= batch_cw : n + pending.front().n <= limit)) {
        s.reqs.push_back(pending.front());
        n += pending.front().n;
        pending_cw -= pending.front().n;
        pending.pop_front();
    }
}

// Enqueue the batch of slot s: inputs, decode, results to pinned staging (or
// straight to device-resident outputs), completion event. Worker thread only,
// without the lock.
void sogrand_server::launch(Slot &s)
{
    bool ok = true, want_soft = false, host_in = false;
    size_t n = 0;
    for (const Request &r : s.reqs) {
        want_soft = want_soft || r.soft;
        host_in = host_in || !r.llr_dev;
        n += r.n;
    }
    if (n > batch_cw) ok = false; // more than the slot's buffers hold: fail the batch
    if (host_in && !s.h_llr &&
        hipHostMalloc((void **)&s.h_llr, batch_cw * sg::kCube * sizeof(double),
                      hipHostMallocDefault) != hipSuccess) {
        s.h_llr = nullptr;
        ok = false;
    }
    size_t off = 0;
    for (const Request &r : s.reqs) {
        const size_t bytes = r.n * sg::kCube * sizeof(double);
        double *dst = s.llr + off * sg::kCube;
        if (r.llr_dev) {
            ok = ok && hipMemcpyAsync(dst, r.llr, bytes, hipMemcpyDeviceToDevice, s.st) ==
                           hipSuccess;
        } else if (ok) {
            std::memcpy(s.h_llr + off * sg::kCube, r.llr, bytes);
            ok = hipMemcpyAsync(dst, s.h_llr + off * sg::kCube, bytes, hipMemcpyHostToDevice,
                                s.st) == hipSuccess;
        }
        off += r.n;
    }
    sg::CubeJob job{};
    job.llr = s.llr;
    job.info = s.info;
    job.soft = want_soft ? s.soft : nullptr;
    job.stats = s.stats;
    job.le_scratch = s.le;
    job.queue = s.queue;
    job.n_cw = n;
    sg::fill_job(job, ctx, p);
    sogrand_params q = p;
    // A batch the plain kernel decodes in one pass over the device runs it:
    // four waves per cube finish a cube sooner, and with many batch slots in
    // flight small plain batches keep every CU busy. Larger batches (requests
    // that piled up while all slots were busy) run the configured kernel.
    // Both decode every codeword identically.
    if (n < plain_below) q.backend = SOGRAND_BACKEND_PLAIN;
    // on a lightly loaded device (this batch and everything still decoding at
    // most lanegroup_max codewords; off by default): the lane-group kernel,
    // whose four lanes per line finish a codeword soonest. Its workgroup holds
    // most of a CU's LDS, so under load it stops other batches from overlapping.
    if (inflight_cw + n <= lanegroup_max) q.backend = SOGRAND_BACKEND_LANEGROUP;
    inflight_cw += n;
    const sg::LaunchPlan lp = sg::plan_launch(ctx, q, n);
    ok = ok && sg::launch_cubes(q, lp, job, s.st, nullptr) == hipSuccess;
    // results: device-resident outputs directly, host outputs via pinned staging
    uint8_t *h_info = s.h_out;
    uint32_t *h_stats = (uint32_t *)(s.h_out + batch_cw * sg::kInfoBytes);
    off = 0;
    for (const Request &r : s.reqs) {
        const bool dev_info = r.info_dev;
        ok = ok && hipMemcpyAsync(dev_info ? (void *)r.info : (void *)(h_info + off * sg::kInfoBytes),
                                  s.info + off * sg::kInfoBytes, r.n * sg::kInfoBytes,
                                  dev_info ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                                  s.st) == hipSuccess;
        if (r.stats) {
            const bool dev = r.stats_dev;
            ok = ok && hipMemcpyAsync(dev ? (void *)r.stats : (void *)(h_stats + off * 6),
                                      s.stats + off * 6, r.n * sizeof(sogrand_cw_stats),
                                      dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                                      s.st) == hipSuccess;
        }
        if (r.soft) {
            const bool dev = r.soft_dev;
            ok = ok && hipMemcpyAsync(dev ? (void *)r.soft : (void *)(s.h_soft + off * sg::kCube),
                                      s.soft + off * sg::kCube, r.n * sg::kCube * sizeof(double),
                                      dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                                      s.st) == hipSuccess;
        }
        off += r.n;
    }
    ok = ok && hipEventRecord(s.done, s.st) == hipSuccess;
    // a launch that failed part-way is finished as failed by the worker (its
    // event may not have been recorded)
    s.launch_ok = ok;
    if (!ok) {
        std::lock_guard<std::mutex> g(mu);
        fatal = true;
    }
}

// Results of a finished batch to the requests' host buffers, then report them.
// Worker thread only; takes the lock only to publish.
void sogrand_server::finish(Slot &s, bool event_ok)
{
    const uint8_t *h_info = s.h_out;
    const uint32_t *h_stats = (const uint32_t *)(s.h_out + batch_cw * sg::kInfoBytes);
    size_t off = 0;
    std::vector<sogrand_completion> out;
    out.reserve(s.reqs.size());
    const double t = now();
    for (const Request &r : s.reqs) {
        if (event_ok) {
            if (!r.info_dev)
                std::memcpy(r.info, h_info + off * sg::kInfoBytes, r.n * sg::kInfoBytes);
            if (r.stats && !r.stats_dev)
                std::memcpy(r.stats, h_stats + off * 6, r.n * sizeof(sogrand_cw_stats));
            if (r.soft && !r.soft_dev)
                std::memcpy(r.soft, s.h_soft + off * sg::kCube, r.n * sg::kCube * sizeof(double));
        }
        const int32_t st = event_ok ? SOGRAND_OK : SOGRAND_EHIP;
        out.push_back(sogrand_completion{r.id, st, (uint32_t)r.n, r.t_submit, s.t_start, t});
        off += r.n;
    }
    size_t n_done = 0;
    for (const Request &r : s.reqs) n_done += r.n;
    inflight_cw -= n_done;
    s.reqs.clear();
    s.busy = false;
    std::lock_guard<std::mutex> g(mu);
    for (const sogrand_completion &c : out) completed.push_back(c);
    finished += out.size();
    if (!event_ok) fatal = true;
}

void sogrand_server::run()
{
    (void)hipSetDevice(ctx->device);
    for (;;) {
        bool progress = false;
        for (Slot &s : slots) {
            if (!s.busy) continue;
            if (!s.launch_ok) {
                // nothing more of it will run usefully: wait for what was
                // enqueued, then report its requests as failed
                (void)hipStreamSynchronize(s.st);
                finish(s, false);
                progress = true;
                continue;
            }
            const hipError_t q = hipEventQuery(s.done);
            if (q == hipErrorNotReady) continue;
            finish(s, q == hipSuccess);
            progress = true;
        }
        for (Slot &s : slots) {
            if (s.busy) continue;
            // optional: while the device is busy enough, requests accumulate
            // and leave together as one larger batch
            if (inflight_cw >= hold_above) break;
            {
                std::lock_guard<std::mutex> g(mu);
                take(s);
            }
            if (s.reqs.empty()) break;
            s.busy = true;
            s.t_start = now();
            launch(s);
            progress = true;
        }
        if (progress) cv_done.notify_all();
        bool busy = false;
        for (const Slot &s : slots) busy = busy || s.busy;
        std::unique_lock<std::mutex> lk(mu);
        if (stop && !busy && pending.empty()) break;
        if (!progress) {
            // poll in-flight batches every 20 us; sleep until new work otherwise
            if (busy)
                cv_work.wait_for(lk, std::chrono::microseconds(20));
            else
                cv_work.wait(lk, [&] { return stop || !pending.empty(); });
        }
    }
}

extern "C" {

void sogrand_server_cfg_default(sogrand_server_cfg *c)
{
    if (!c) return;
    std::memset(c, 0, sizeof *c);
    c->hbm_budget = (size_t)4 << 30;
    c->max_inflight = 16;
    c->max_batch_cw = 0;
    c->max_pending_cw = 1u << 20;
    c->with_soft = 0;
    c->plain_below_cw = 0;
    c->hold_above_cw = 0;
    c->lanegroup_max_cw = 0;
    c->small_batch_cw = 256;
    c->backlog_cw = 0;
}

int sogrand_server_create(sogrand_server **out, sogrand_ctx *ctx, const sogrand_server_cfg *cfg)
{
    if (!out || !ctx) return SOGRAND_EINVAL;
    *out = nullptr;
    sogrand_server_cfg c;
    if (cfg)
        c = *cfg;
    else
        sogrand_server_cfg_default(&c);
    if (c.max_inflight < 1 || c.max_inflight > 32) return SOGRAND_EINVAL;
    sogrand_server *s = new (std::nothrow) sogrand_server;
    if (!s) return SOGRAND_ENOMEM;
    s->ctx = ctx;
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        s->p = ctx->p;
    }
    s->cfg = c;
    // per-codeword bytes of a slot: LLRs, L_E scratch (up to
    // kPersistentSlotScratch cubes per codeword for the persistent kernel),
    // optional soft output, bits, statistics
    const size_t per_cw = (1 + sg::kPersistentSlotScratch) * sg::kCube * sizeof(double) +
                          (c.with_soft ? sg::kCube * sizeof(double) : 0) + sg::kInfoBytes +
                          sizeof(sogrand_cw_stats);
    size_t b = c.hbm_budget / c.max_inflight / per_cw;
    if (b > 65536) b = 65536;
    if (c.max_batch_cw && b > c.max_batch_cw) b = c.max_batch_cw;
    if (b < 1) {
        delete s;
        return SOGRAND_EINVAL; // budget too small for one codeword per slot
    }
    s->batch_cw = b;
    s->plain_below = c.plain_below_cw ? c.plain_below_cw : (size_t)3 * ctx->cu_count;
    s->hold_above = c.hold_above_cw ? c.hold_above_cw : SIZE_MAX;
    // the lane-group kernel has the final decoder's FP64 arithmetic only
    const bool lg_ok = s->p.precision == SOGRAND_FP64 && s->p.variant == SOGRAND_VARIANT_FINAL;
    s->lanegroup_max = lg_ok ? (size_t)c.lanegroup_max_cw : 0;
    s->small_batch = c.small_batch_cw;
    s->backlog = c.backlog_cw ? (size_t)c.backlog_cw : (size_t)8 * c.small_batch_cw;
    if (hipSetDevice(ctx->device) != hipSuccess) {
        delete s;
        return SOGRAND_EHIP;
    }
    s->slots.resize(c.max_inflight);
    for (Slot &sl : s->slots) {
        const size_t le_n = sg::max_le_slots(b) * sg::kCube;
        const size_t bytes_d = (b * sg::kCube + le_n + (c.with_soft ? b * sg::kCube : 0)) *
                               sizeof(double);
        const size_t bytes = bytes_d + b * 6 * sizeof(uint32_t) + 256 + b * sg::kInfoBytes;
        if (hipMalloc(&sl.base, bytes) != hipSuccess) {
            sl.base = nullptr;
            sogrand_server_destroy(s);
            return SOGRAND_ENOMEM;
        }
        char *q = (char *)sl.base;
        sl.llr = (double *)q;
        sl.le = sl.llr + b * sg::kCube;
        sl.soft = c.with_soft ? sl.le + le_n : nullptr;
        q += bytes_d;
        sl.stats = (uint32_t *)q;
        q += b * 6 * sizeof(uint32_t);
        sl.queue = (uint32_t *)q;
        q += 256;
        sl.info = (uint8_t *)q;
        const size_t out_bytes = b * (sg::kInfoBytes + 6 * sizeof(uint32_t));
        if (hipHostMalloc((void **)&sl.h_out, out_bytes, hipHostMallocDefault) != hipSuccess ||
            (c.with_soft && hipHostMalloc((void **)&sl.h_soft, b * sg::kCube * sizeof(double),
                                          hipHostMallocDefault) != hipSuccess)) {
            sogrand_server_destroy(s);
            return SOGRAND_ENOMEM;
        }
        if (hipStreamCreateWithFlags(&sl.st, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess) {
            sogrand_server_destroy(s);
            return SOGRAND_EHIP;
        }
    }
    s->t0 = std::chrono::steady_clock::now();
    s->worker = std::thread([s] { s->run(); });
    *out = s;
    return SOGRAND_OK;
}

int sogrand_server_batch_cw(const sogrand_server *s, size_t *batch_cw)
{
    if (!s || !batch_cw) return SOGRAND_EINVAL;
    *batch_cw = s->batch_cw;
    return SOGRAND_OK;
}

int sogrand_server_submit(sogrand_server *s, uint64_t id, const double *llr, size_t n_cw,
                          uint8_t *info_bits, double *soft_out, sogrand_cw_stats *stats)
{
    if (!s || !llr || !info_bits || n_cw == 0) return SOGRAND_EINVAL;
    if (n_cw > s->batch_cw) return SOGRAND_EINVAL; // larger than one batch slot
    if (soft_out && !s->cfg.with_soft) return SOGRAND_EINVAL;
    Request r{id,
              llr,
              n_cw,
              info_bits,
              soft_out,
              stats,
              0.0,
              sg::is_device_ptr(llr),
              sg::is_device_ptr(info_bits),
              soft_out && sg::is_device_ptr(soft_out),
              stats && sg::is_device_ptr(stats)};
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (s->fatal) return SOGRAND_EHIP;
        if (s->pending_cw + n_cw > s->cfg.max_pending_cw) return SOGRAND_EBUSY;
        r.t_submit = s->now();
        s->pending.push_back(r);
        s->pending_cw += n_cw;
        ++s->submitted;
    }
    s->cv_work.notify_one();
    return SOGRAND_OK;
}

int sogrand_server_poll(sogrand_server *s, sogrand_completion *out, size_t max, size_t *n_out,
                        int wait_ms)
{
    if (!s || !n_out || (!out && max)) return SOGRAND_EINVAL;
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->completed.empty() && wait_ms > 0)
        s->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms),
                            [&] { return !s->completed.empty() || s->fatal; });
    size_t k = 0;
    while (k < max && !s->completed.empty()) {
        out[k++] = s->completed.front();
        s->completed.pop_front();
    }
    *n_out = k;
    return s->fatal ? SOGRAND_EHIP : SOGRAND_OK;
}

int sogrand_server_drain(sogrand_server *s)
{
    if (!s) return SOGRAND_EINVAL;
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->finished == s->submitted || s->fatal; });
    return s->fatal ? SOGRAND_EHIP : SOGRAND_OK;
}

double sogrand_server_clock(const sogrand_server *s) { return s ? s->now() : 0.0; }

void sogrand_server_destroy(sogrand_server *s)
{
    if (!s) return;
    if (s->worker.joinable()) {
        {
            std::lock_guard<std::mutex> g(s->mu);
            s->stop = true;
        }
        s->cv_work.notify_all();
        s->worker.join();
    }
    (void)hipSetDevice(s->ctx->device);
    for (Slot &sl : s->slots) {
        if (sl.st) (void)hipStreamSynchronize(sl.st);
        if (sl.base) (void)hipFree(sl.base);
        if (sl.h_out) (void)hipHostFree(sl.h_out);
        if (sl.h_soft) (void)hipHostFree(sl.h_soft);
        if (sl.h_llr) (void)hipHostFree(sl.h_llr);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.st) (void)hipStreamDestroy(sl.st);
    }
    delete s;
}

} // extern "C"
