// ssa_kernels.hip -- ahead-of-time gfx950 kernels of the batched SSA.
//
// The lane engine (ssa_lane.h) runs here with a data-driven network policy,
// GenericNet<MS, MR>: any network with S <= MS species and R <= MR reactions.
// Counts live in VGPRs as exact FP64 integers (x[MS]); reaction data is read
// at wave-uniform addresses (scalar cache -> SGPRs), so x[fac_s] compiles to
// an s_set_gpr_idx-indexed register move; the net stoichiometry table is
// staged in LDS, where each lane gathers the row of its chosen reaction.
#include <hip/hip_runtime.h>

#include "rebop_rng.h"
#include "ssa_lane.h"
#include "ssa_layout.h"

#pragma clang fp contract(off)

namespace rebop {

template <int MS_, int MR_, bool CAP>
struct GenericNet {
  static constexpr int MS = MS_;
  static constexpr int MR = MR_;
  static constexpr bool kCap = CAP;
  // f64 counts: a run leaving |x| <= 2^52 is rerun on the int64 LDS kernel
  static constexpr uint32_t kRangeFlag = REBOP_FLAG_EXACT;
  static constexpr uint32_t kStopFlag = 0u; // a run that is rerun finishes first
  rebop_dev_net net;
  const double *lds_diff; // [R][S]
  double kr[MR];          // this trajectory's rate constants

  __device__ __forceinline__ double grid_time(const rebop_dev_run &run, uint64_t i) const {
    return table_time(run, i);
  }

  // the network's constants, or (parameter sweeps) this trajectory's row
  __device__ __forceinline__ void start(const rebop_dev_run &run, uint64_t j) {
    const double *src = run.rates ? run.rates + (j / run.rates_div) * net.R : net.k;
#pragma unroll
    for (int r = 0; r < MR; ++r) kr[r] = r < (int)net.R ? src[r] : 0.0;
  }
  __device__ __forceinline__ uint32_t S() const { return net.S; }
  __device__ __forceinline__ void load_x0(double (&x)[MS], const int64_t *src) const {
    reg_load_x0(x, src, net.S);
  }
  __device__ __forceinline__ void load_state(double (&x)[MS], const double *st, uint64_t stride) const {
    reg_load_state(x, st, stride, net.S);
  }
  __device__ __forceinline__ void store_state(const double (&x)[MS], double *st, uint64_t stride) const {
    reg_store_state(x, st, stride, net.S);
  }
  __device__ __forceinline__ void write_raw(const double (&x)[MS], int64_t *dst, uint64_t stride) const {
    reg_write_raw(x, dst, stride, net.S);
  }
  __device__ __forceinline__ void write_counts(const double (&x)[MS], int64_t *dst, uint64_t stride) const {
    reg_write_counts(x, dst, stride, net.S);
  }
  __device__ __forceinline__ bool in_range(const double (&x)[MS]) const { return reg_in_range(x, net.S); }

  // make_cumrates (src/gillespie.rs:357-364) over Rate::rate (:71-90)
  __device__ __forceinline__ double props(double (&x)[MS], double (&cum)[MR]) const {
    double acc = 0.0;
#pragma unroll
    for (int r = 0; r < MR; ++r) {
      if (r < (int)net.R) {
        double a = kr[r];
        const uint32_t fb = net.fac_begin[r];
        const uint32_t fe = net.fac_begin[r + 1];
        for (uint32_t f = fb; f < fe; ++f) a = a * (x[net.fac_s[f]] - net.fac_c[f]);
        acc = acc + a;
        cum[r] = acc;
      }
    }
    return acc;
  }

  // choose_cumrate_sum (src/gillespie.rs:402-407)
  __device__ __forceinline__ uint32_t choose(double (&)[MS], double (&cum)[MR], double chosen) const {
    uint32_t k = 0;
#pragma unroll
    for (int r = 0; r < MR; ++r)
      if (r < (int)net.R) k += (cum[r] < chosen) ? 1u : 0u;
    return k;
  }

  // Jump::affect (src/gillespie.rs:142-152) with the dense row in LDS
  __device__ __forceinline__ void affect(double (&x)[MS], uint32_t k) const {
    const double *drow = lds_diff + (uint64_t)k * net.S;
#pragma unroll
    for (int s = 0; s < MS; ++s)
      if (s < (int)net.S) x[s] = x[s] + drow[s];
  }
};

template <int MS, int MR, bool CAP>
__global__ void __launch_bounds__(256) ssa_reg_kernel(rebop_dev_net net, rebop_dev_run run) {
  extern __shared__ double lds_diff[]; // [R][S]
  __shared__ unsigned long long blk_ctr[4];
  for (uint32_t q = threadIdx.x; q < net.R * net.S; q += blockDim.x) lds_diff[q] = net.diff[q];
  if (threadIdx.x < 4) blk_ctr[threadIdx.x] = 0ull;
  __syncthreads();
  GenericNet<MS, MR, CAP> pol;
  pol.net = net;
  pol.lds_diff = lds_diff;
  const uint64_t gtid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t nthreads = (uint64_t)gridDim.x * blockDim.x;
  const LaneTotals acc = run.mode == REBOP_MODE_EVENTS ? lane_events(pol, run, gtid, nthreads)
                                                       : lane_loop(pol, run, gtid, nthreads);
  flush_totals(acc, run, blk_ctr);
}

// The sample times tmax * i / nb_steps (src/pyo3_gillespie.rs:202: product,
// then quotient, as grid_time computes them), or tmax for a target run.
__global__ void grid_kernel(double *grid, double tmax, uint64_t nb, uint64_t n, int is_grid) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) grid[i] = is_grid ? grid_time(tmax, i, nb) : tmax;
}

// Clears the launch counters with a kernel (not hipMemsetAsync), so that the
// sequence stays valid inside a captured hipGraph.
__global__ void reset_counters_kernel(unsigned long long *counters) {
  if (threadIdx.x < REBOP_CTR_COUNT) counters[threadIdx.x] = 0ull;
}

// Clears n 64-bit words (per-job event counters) with a kernel, for the same reason.
__global__ void zero_u64_kernel(unsigned long long *p, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0ull;
}

} // namespace rebop

// ---------------------------------------------------------------------------
// launchers (C++ linkage, used by rebop_gpu.cpp)

namespace rebop {

struct RegShape {
  int ms, mr;
};

static const RegShape kRegShapes[] = {{4, 8}, {8, 16}, {16, 32}};

int reg_shape_count() { return (int)(sizeof(kRegShapes) / sizeof(kRegShapes[0])); }
int reg_shape_ms(int c) { return kRegShapes[c].ms; }
int reg_shape_mr(int c) { return kRegShapes[c].mr; }

static const void *reg_kernel_ptr(int c) {
  switch (c) {
  case 0: return (const void *)&ssa_reg_kernel<4, 8, false>;
  case 1: return (const void *)&ssa_reg_kernel<8, 16, false>;
  case 2: return (const void *)&ssa_reg_kernel<16, 32, false>;
  default: return nullptr;
  }
}

hipError_t reg_occupancy(int c, int threads, size_t lds, int *blocks_per_cu) {
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, reg_kernel_ptr(c), threads,
                                                      lds);
}

hipError_t launch_reset_counters(unsigned long long *counters, hipStream_t stream) {
  hipLaunchKernelGGL(reset_counters_kernel, dim3(1), dim3(64), 0, stream, counters);
  return hipGetLastError();
}

hipError_t launch_zero_u64(unsigned long long *p, uint64_t n, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  uint64_t blocks = (n + 255u) / 256u;
  if (blocks > 1024u) blocks = 1024u;
  hipLaunchKernelGGL(zero_u64_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, p, n);
  return hipGetLastError();
}

// out[k] holds the f64 bits of an exact integer count (|x| <= 2^52 unless
// the run reports a range error): replace it by the int64 value.
__global__ void raw_counts_kernel(int64_t *out, uint64_t n, int pairs) {
  if (!pairs) { // an output that is only 8-byte aligned
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride)
      out[k] = (int64_t)rebop_f64_from_bits((uint64_t)out[k]);
    return;
  }
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x * 2u;
  for (uint64_t k = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * 2u; k < n; k += stride) {
    if (k + 1u < n) {
      longlong2 *p = (longlong2 *)(out + k); // 16-byte aligned: out is, and k is even
      const longlong2 v = *p;
      longlong2 r;
      r.x = (long long)rebop_f64_from_bits((uint64_t)v.x);
      r.y = (long long)rebop_f64_from_bits((uint64_t)v.y);
      *p = r;
    } else {
      out[k] = (int64_t)rebop_f64_from_bits((uint64_t)out[k]);
    }
  }
}

hipError_t launch_raw_counts(int64_t *out, uint64_t n, int cus, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  uint64_t blocks = (n / 2u + 255u) / 256u;
  const uint64_t cap = (uint64_t)(cus > 0 ? cus : 256) * 16u;
  if (blocks > cap) blocks = cap;
  if (blocks == 0) blocks = 1;
  const int pairs = ((uintptr_t)out & 15u) == 0u ? 1 : 0;
  hipLaunchKernelGGL(raw_counts_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, out, n, pairs);
  return hipGetLastError();
}

hipError_t launch_grid(double *grid, const rebop_dev_run &run, hipStream_t stream) {
  const uint64_t n = run.mode == REBOP_MODE_GRID ? run.nb_steps + 1u : 1u;
  const unsigned blocks = (unsigned)((n + 255u) / 256u);
  hipLaunchKernelGGL(grid_kernel, dim3(blocks), dim3(256), 0, stream, grid, run.tmax, run.nb_steps,
                     n, run.mode == REBOP_MODE_GRID ? 1 : 0);
  return hipGetLastError();
}

// the variant with the draw-cap checks when the run has a cap
hipError_t launch_reg(int c, const rebop_dev_net &net, const rebop_dev_run &run, int blocks,
                      int threads, size_t lds, hipStream_t stream) {
  const bool cap = run.draw_cap != 0;
  switch (c * 2 + (cap ? 1 : 0)) {
  case 0: hipLaunchKernelGGL((ssa_reg_kernel<4, 8, false>), dim3(blocks), dim3(threads), lds, stream, net, run); break;
  case 1: hipLaunchKernelGGL((ssa_reg_kernel<4, 8, true>), dim3(blocks), dim3(threads), lds, stream, net, run); break;
  case 2: hipLaunchKernelGGL((ssa_reg_kernel<8, 16, false>), dim3(blocks), dim3(threads), lds, stream, net, run); break;
  case 3: hipLaunchKernelGGL((ssa_reg_kernel<8, 16, true>), dim3(blocks), dim3(threads), lds, stream, net, run); break;
  case 4: hipLaunchKernelGGL((ssa_reg_kernel<16, 32, false>), dim3(blocks), dim3(threads), lds, stream, net, run); break;
  case 5: hipLaunchKernelGGL((ssa_reg_kernel<16, 32, true>), dim3(blocks), dim3(threads), lds, stream, net, run); break;
  default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

} // namespace rebop
