// noalloc.hip -- rebop_run_device_noalloc: the grid run of rebop_run_device
// for callers that cannot allocate device memory (code being captured into a
// HIP graph, serving loops, harnesses that forbid hipMalloc).
//
// The launch counters, the sample-time table and the LDS kernel's work tables
// live in one fixed device array per GPU, `scratch` below, instead of the
// handle's buffers.  Its device address is taken once per device with
// hipGetSymbolAddress, and only that address is handed to the launch path
// (the array's name is a host-side stand-in, not a device address).  Every
// call on a GPU shares its array, so calls that overlap on one GPU race:
// issue them on one stream, one after another.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/rebop_gpu.h"
#include "scratch.h"

// one GPU's scratch region, ~0.8 MB (each GPU the library runs on keeps its
// copy until the process ends)
static __device__ __align__(256) unsigned char scratch[rebop::kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char *g_scratch[kMaxDevices] = {};

// The device address of the current GPU's scratch array (looked up once per device).
int scratch_address(unsigned char **out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return rebop::report_error(REBOP_ERR_DEVICE, "no HIP device available");
  if (dev < 0 || dev >= kMaxDevices)
    return rebop::report_error(REBOP_ERR_DEVICE, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[dev]) {
    void *ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return rebop::report_error(
          REBOP_ERR_DEVICE,
          (std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e)).c_str());
    g_scratch[dev] = (unsigned char *)ptr;
  }
  *out = g_scratch[dev];
  return REBOP_OK;
}

} // namespace

extern "C" int rebop_run_device_noalloc(const rebop_gillespie *g, double tmax, uint64_t nb_steps,
                                        uint64_t traj_begin, uint64_t n_traj,
                                        const int64_t *d_x0, int64_t *d_out, int64_t *d_final,
                                        int kernel, void *stream, rebop_stats *stats) {
  if (nb_steps + 1u > rebop::kScratchMaxSamples || nb_steps >= (1ull << 40))
    return rebop::report_error(REBOP_ERR_VALUE,
                               ("nb_steps = " + std::to_string(nb_steps) +
                                " exceeds the no-allocation limit of " +
                                std::to_string(rebop::kScratchMaxSamples - 1u))
                                   .c_str());
  unsigned char *base = nullptr;
  if (int rc = scratch_address(&base)) return rc;
  const rebop::Scratch sc{base};
  return rebop::run_device_scratch(g, tmax, nb_steps, traj_begin, n_traj, d_x0, d_out, d_final,
                                   kernel, stream, stats, &sc);
}
