// Internal: the decoder context behind the opaque sogrand_ctx handle, shared
// by the C ABI (sg_api.cpp) and the serving loop (sg_serve.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>

#include "sg_launch.h"
#include "sogrand.h"

namespace sg {

// Grow-only device buffer.
template <class T> struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    hipError_t reserve(size_t want)
    {
        if (want <= n) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        hipError_t e = hipMalloc(&p, sizeof(T) * want);
        if (e != hipSuccess) {
            p = nullptr;
            return e;
        }
        n = want;
        return hipSuccess;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

// True for device (or managed) memory, false for host memory of any kind.
bool is_device_ptr(const void *ptr);

// Guess budget as the kernels take it: 0 means 0x7fffffff, as the reference
// maps it (cubic_decoder_final.c:14, :607).
inline uint64_t effective_tmax(uint64_t t) { return t == 0 ? 0x7fffffffull : t; }

} // namespace sg

// A captured decode launch sequence and what it was captured for.
struct GraphKey {
    const void *llr, *info, *soft, *stats, *le, *queue;
    size_t n;
    hipStream_t st;
    sogrand_params p;
};

// One of the two alternating stages of a decode of host buffers: its own
// stream, staging buffers, L_E scratch, work counter and timing events,
// so that one chunk's input copy and kernel overlap the previous chunk's kernel.
struct HostStage {
    hipStream_t st = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    sg::DevBuf<double> llr, soft, le;
    sg::DevBuf<uint8_t> info;
    sg::DevBuf<uint32_t> stats, queue;
    size_t off = 0, n = 0; // the chunk in flight
    bool busy = false;
};

struct sogrand_ctx {
    int device = 0;
    sogrand_params p{};
    hipStream_t own_stream = nullptr;
    uint16_t *d_sched = nullptr;
    sg::DevBuf<double> le_scratch, llr_stage, soft_stage, sigma_stage;
    sg::DevBuf<uint8_t> info_stage;
    sg::DevBuf<uint32_t> stats_stage, queue;
    sg::DevBuf<uint32_t> nguess_stage;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0.0;
    uint32_t last_launches = 0;
    int cu_count = 256;
    hipGraphExec_t graph = nullptr;
    GraphKey graph_key{};
    uint32_t graph_replays = 0; // replays of the cached graph (tests)
    HostStage hs[2];
    std::mutex mu;
};

namespace sg {

// Grid and L_E scratch of one decode launch of n codewords.
struct LaunchPlan {
    bool persistent;
    bool lanegroup;
    int blocks;
    int slots;       // persistent: cube slots per wave
    size_t le_slots; // 4096-f64 L_E slots the launch needs
};
LaunchPlan plan_launch(const sogrand_ctx *ctx, const sogrand_params &p, size_t n);
// Upper bound of plan_launch(...).le_slots for any parameters and n codewords.
size_t max_le_slots(size_t n);
// Decoder constants of a job from the parameters.
void fill_job(CubeJob &job, const sogrand_ctx *ctx, const sogrand_params &p);
// Enqueue one decode (queue reset + kernel) on st; ev_start (optional) is
// recorded right before the kernel.
hipError_t launch_cubes(const sogrand_params &p, const LaunchPlan &lp, const CubeJob &job,
                        hipStream_t st, hipEvent_t ev_start);

} // namespace sg
