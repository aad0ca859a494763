// sogrand_decode_cubic_noalloc: a decode that allocates nothing, for callers
// that cannot call hipMalloc (code under HIP graph capture, tight serving
// loops, harnesses that forbid allocation). It runs the plain kernel of
// sogrand_decode_cubic on caller-owned device buffers; the kernel's device
// scratch (one L_E cube per workgroup) lives in a file-scope static device
// array, `scratch`, of SOGRAND_NOALLOC_SCRATCH_BYTES per GPU, held by every GPU
// the library runs on until the process ends.
//
// The array's name is a host-side stand-in in host code: its device address
// is taken once per device with hipGetSymbolAddress (noalloc_scratch_init,
// called by sogrand_ctx_create with the device current) and only that address
// is passed to kernels and copies. Every call on a GPU shares that GPU's array
// and leaves its contents behind: calls that overlap on one GPU race (the
// context's own mutex serialises calls on one context only).
#include <hip/hip_runtime.h>

#include <mutex>

#include "sg_common.h"
#include "sg_ctx.h"
#include "sg_launch.h"
#include "sogrand.h"

namespace sg {

namespace {

// The array holds the L_E cubes of the plain kernel's workgroups (one each; the
// kernel loops over codewords, so fewer workgroups than codewords is fine).
constexpr size_t kCubeBytes = (size_t)kCube * sizeof(double);
constexpr size_t kNoAllocCubes = SOGRAND_NOALLOC_SCRATCH_BYTES / kCubeBytes;
static_assert(kNoAllocCubes >= 768, "one scratch cube per resident plain workgroup");

static __device__ __align__(256) unsigned char scratch[SOGRAND_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char *g_ptr[kMaxDevices] = {}; // device address of `scratch`, per device

} // namespace

// Take the device address of `scratch` on `device` (which must be current),
// once per device. Called by sogrand_ctx_create.
hipError_t noalloc_scratch_init(int device)
{
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(g_mu);
    if (g_ptr[device]) return hipSuccess;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return e != hipSuccess ? e : hipErrorInvalidSymbol;
    g_ptr[device] = static_cast<unsigned char *>(p);
    return hipSuccess;
}

static unsigned char *scratch_of(int device)
{
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> g(g_mu);
    return g_ptr[device];
}

} // namespace sg

extern "C" int sogrand_decode_cubic_noalloc(sogrand_ctx *ctx, const double *llr, size_t n_cw,
                                            uint8_t *info_bits, double *soft_out,
                                            sogrand_cw_stats *stats, void *hip_stream)
{
    if (!ctx || (!llr && n_cw) || (!info_bits && n_cw)) return SOGRAND_EINVAL;
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->last_ms = 0.0;
    ctx->last_launches = 0;
    if (n_cw == 0) return SOGRAND_OK;
    // device buffers only: host buffers would need staging memory
    if (!sg::is_device_ptr(llr) || !sg::is_device_ptr(info_bits) ||
        (soft_out && !sg::is_device_ptr(soft_out)) || (stats && !sg::is_device_ptr(stats)))
        return SOGRAND_EINVAL;
    if (hipSetDevice(ctx->device) != hipSuccess) return SOGRAND_EHIP;
    unsigned char *base = sg::scratch_of(ctx->device);
    if (!base) return SOGRAND_EHIP;
    hipStream_t st = hip_stream ? (hipStream_t)hip_stream : ctx->own_stream;

    sg::CubeJob job{};
    job.llr = llr;
    job.info = info_bits;
    job.soft = soft_out;
    job.stats = reinterpret_cast<uint32_t *>(stats);
    job.le_scratch = reinterpret_cast<double *>(base);
    job.queue = nullptr; // the plain kernel has no work counter
    job.n_cw = n_cw;
    sg::fill_job(job, ctx, ctx->p);
    size_t grid = sg::plain_scratch_slots(n_cw);
    if (grid > sg::kNoAllocCubes) grid = sg::kNoAllocCubes;

    (void)hipEventRecord(ctx->ev0, st);
    if (sg::launch_plain(job, (int)grid, st) != hipSuccess) return SOGRAND_EHIP;
    (void)hipEventRecord(ctx->ev1, st);
    if (hipStreamSynchronize(st) != hipSuccess) return SOGRAND_EHIP;
    float t = 0.f;
    (void)hipEventElapsedTime(&t, ctx->ev0, ctx->ev1);
    ctx->last_ms = t;
    ctx->last_launches = 1;
    return SOGRAND_OK;
}
