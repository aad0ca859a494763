// Host-side launch interface between the C ABI (sg_api.cpp) and the HIP
// kernels (sg_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace sg {

// Arithmetic policy of the line decoder (sg_device.h): the reference's
// cubic_decoder_final.c in double (PrecF64) or with single-precision path
// metrics (PrecF32), or its earlier cubic_decoder.c in double (PrecF64Legacy).
enum { kArithF64 = 0, kArithF32 = 1, kArithF64Legacy = 2 };

struct CubeJob {
    const double *llr;        // n_cw x 4096 (device)
    uint8_t *info;            // n_cw x 64 (device)
    double *soft;             // n_cw x 4096 or null (device)
    uint32_t *stats;          // n_cw x 6 u32 {half, conv, NG lo, NG hi, NGp lo, NGp hi} or null
    double *le_scratch;       // 4096-f64 cubes: one per plain workgroup, three
                              // per persistent wave slot (kPersistentSlotScratch)
    uint32_t *queue;          // persistent: work counter (device), zeroed by the launcher
    size_t n_cw;
    const uint16_t *sched;
    uint64_t tmax;
    double thres;
    double alpha;
    uint32_t L;
    uint32_t Imax;
    int arith;                // kArithF64 / kArithF32 / kArithF64Legacy
    double eps;               // EPSILON of the confidence test, P_notinlist floor and clamps
    double eps_log;           // floor inside the soft output's logs
};

// Whether a guess budget can stop a search. A line tests at most 65536 patterns
// (the hard decision and 65535 others), so the reference's guard T < Tmax,
// taken before each weight level with T <= 65535 (cubic_decoder_final.c:652),
// can only fail for Tmax <= 65535; larger budgets run the unbounded kernels.
inline bool tmax_binds(uint64_t tmax) { return tmax < 65536u; }

// Plain backend: one 256-lane workgroup per cube, one lane per line.
hipError_t launch_plain(const CubeJob &job, int grid, hipStream_t st);
size_t plain_scratch_slots(size_t n_cw);

// The (16,8)^2 square code: sixteen squares per 256-lane workgroup, 256 LLRs and
// 8 info bytes per square (llr/info/soft of the job in those units); the job's
// scratch holds one 4096-double cube per workgroup.
hipError_t launch_square(const CubeJob &job, int grid, hipStream_t st);
size_t square_blocks(size_t n_squares);

// Lane-group backend: one 1024-lane workgroup per cube (cube in LDS, four lanes
// per line), at most one resident workgroup per CU; FP64, final variant.
hipError_t launch_lanegroup(const CubeJob &job, int grid, hipStream_t st);
size_t lanegroup_blocks(size_t n_cw, int cu_count);

// Persistent backend: waves pop cubes from a device counter; each wave keeps
// `slots` cubes in flight and hands lines to lanes as they free up.
hipError_t launch_persistent(const CubeJob &job, int waves_per_block, int blocks, int slots,
                             uint32_t lane_batch, uint32_t step_len, hipStream_t st);
// Device address of the no-allocation entry point's static scratch on
// `device` (current), taken once per device (sg_noalloc.hip).
hipError_t noalloc_scratch_init(int device);
// Resident workgroups per CU of the persistent kernel (occupancy API).
int persistent_blocks_per_cu(int slots, bool finite, int arith);
constexpr int kPersistentWaves = 4;
// 4096-f64 scratch cubes per persistent wave slot: two L_E cubes used in turn
// and the current phase's line inputs (sg_persistent.hip)
constexpr int kPersistentSlotScratch = 3;

// Line-level SISO (SOGRAND_bitSO on many lines).
hipError_t launch_lines(const double *llr, size_t n_lines, double *la, double *le,
                        uint32_t *nguess, const uint16_t *sched, uint64_t tmax,
                        double thres, uint32_t L, double eps, double eps_log, int arith,
                        hipStream_t st);

// Synthetic LLR generation on the device.
hipError_t launch_gen(uint64_t seed, uint64_t cw_offset, size_t n_cw, const double *sigma,
                      size_t n_sigma, double *llr, hipStream_t st);

// Fill kernel (used instead of hipMemsetAsync so that callers may capture it).
hipError_t launch_fill_u32(uint32_t *p, uint32_t v, size_t n, hipStream_t st);

} // namespace sg
