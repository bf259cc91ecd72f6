// scratch.h -- layout of the fixed device scratch region that
// rebop_run_device_noalloc (noalloc.hip) hands to the launch path in
// rebop_gpu.cpp instead of per-handle buffers.
#ifndef REBOP_SCRATCH_H
#define REBOP_SCRATCH_H

#include <stddef.h>
#include <stdint.h>

#include "../../include/rebop_gpu.h"
#include "ssa_layout.h"

namespace rebop {

// Device address of one GPU's scratch region (from hipGetSymbolAddress).
struct Scratch {
  unsigned char *base;
};

constexpr size_t scratch_align(size_t b) { return (b + 255u) & ~(size_t)255u; }

// launch counters (REBOP_CTR_COUNT 64-bit words)
constexpr size_t kScratchCounters = 0;
// the LDS kernel's job record (one network)
constexpr size_t kScratchJob = 256;
constexpr size_t kScratchJobBytes = 768;
// sample times tmax * i / nb_steps, i = 0 .. nb_steps
constexpr size_t kScratchMaxSamples = 65537; // nb_steps <= 65536
constexpr size_t kScratchGrid = kScratchJob + kScratchJobBytes;
// the LDS kernel's work items (64 trajectories each)
constexpr size_t kScratchMaxItems = 16384; // 2^20 trajectories
constexpr size_t kScratchItems = scratch_align(kScratchGrid + sizeof(double) * kScratchMaxSamples);
constexpr size_t kScratchItemsBytes = sizeof(rebop_dev_item) * kScratchMaxItems;
constexpr size_t kScratchBytes = kScratchItems + kScratchItemsBytes;

static_assert(sizeof(unsigned long long) * REBOP_CTR_COUNT <= kScratchJob, "counters fit");
static_assert(sizeof(rebop_dev_job) <= kScratchJobBytes, "job record fits");

// rebop_run_device with `sc` in place of the handle's buffers (rebop_gpu.cpp).
int run_device_scratch(const rebop_gillespie *g, double tmax, uint64_t nb_steps,
                       uint64_t traj_begin, uint64_t n_traj, const int64_t *d_x0, int64_t *d_out,
                       int64_t *d_final, int kernel, void *stream, rebop_stats *stats,
                       const Scratch *sc);

// Sets this thread's last-error message and returns `code` (rebop_gpu.cpp).
int report_error(int code, const char *msg);

} // namespace rebop

#endif
