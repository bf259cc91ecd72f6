// ssa_kernels.hip -- gfx950 kernels of the batched direct-method SSA.
//
// One trajectory per lane.  Each lane runs rebop's grid loop
// (src/pyo3_gillespie.rs:200-208) around Gillespie::advance_until
// (src/gillespie.rs:315-344) as a state machine whose every iteration is one
// exponential draw:
//
//   tot <= 0 or NaN  -> absorbing: every remaining sample gets the current
//                       counts, no draw (gillespie.rs:323-326)
//   t + E/tot > T_i  -> overshoot: t = T_i, sample i, the draw is discarded
//                       (gillespie.rs:327-331); counts are unchanged, so the
//                       cumulative propensities are reused for T_{i+1}
//   otherwise        -> event: chosen = tot*U, k = #{r : cum_r < chosen}
//                       (choose_cumrate_sum, gillespie.rs:402-407), apply
//                       reaction k (Jump::affect, :142-152), recompute the
//                       propensities in the reference's order (make_cumrates
//                       :357-364 over Rate::rate :71-90)
//
// Counts live in VGPRs as exact FP64 integers (x[MS]); reaction data is read
// at wave-uniform addresses (scalar cache -> SGPRs), so x[fac_s] compiles to
// an s_set_gpr_idx-indexed register move; the net stoichiometry table is
// staged in LDS, where each lane gathers the row of its chosen reaction.
//
// The persistent variant keeps lanes alive: a lane whose trajectory is done
// takes the next trajectory id from a device-wide counter, one atomic per
// wave per refill (ballot + mbcnt), so short trajectories do not leave
// lanes idle while the wave's longest trajectory finishes.
#include <hip/hip_runtime.h>

#include "rebop_rng.h"
#include "ssa_layout.h"

#pragma clang fp contract(off)

namespace rebop {

__device__ __forceinline__ double grid_time(const rebop_dev_run &run, uint64_t i) {
  if (run.mode == REBOP_MODE_TARGET) return run.tmax;
  // src/pyo3_gillespie.rs:202: tmax * i as f64 / nb_steps as f64
  return run.tmax * (double)i / (double)run.nb_steps;
}

template <int MS, int MR>
__global__ void __launch_bounds__(256) ssa_reg_kernel(rebop_dev_net net, rebop_dev_run run) {
  extern __shared__ double lds_diff[]; // [R][S]
  __shared__ unsigned long long blk_ctr[4];

  const uint32_t S = net.S;
  const uint32_t R = net.R;
  for (uint32_t q = threadIdx.x; q < R * S; q += blockDim.x) lds_diff[q] = net.diff[q];
  if (threadIdx.x < 4) blk_ctr[threadIdx.x] = 0ull;
  __syncthreads();

  const uint64_t n = run.n_traj;
  const uint64_t nb = run.mode == REBOP_MODE_TARGET ? 0ull : run.nb_steps;
  const uint64_t total_threads = (uint64_t)gridDim.x * blockDim.x;

  double x[MS];
  double cum[MR];
  double tot = 0.0, t = 0.0, T = 0.0;
  uint64_t i = 0, draw = 0, draw_start = 0, draw_mark = 0, jl = 0, traj = 0;
  uint32_t ov = 0, flags = 0;
  uint64_t ev_acc = 0, dr_acc = 0, tr_acc = 0;

#pragma unroll
  for (int s = 0; s < MS; ++s) x[s] = 0.0;
#pragma unroll
  for (int r = 0; r < MR; ++r) cum[r] = 0.0;

  // make_cumrates over Rate::rate, in the reference's order
  auto props = [&]() {
    double acc = 0.0;
#pragma unroll
    for (int r = 0; r < MR; ++r) {
      if (r < (int)R) {
        double a = net.k[r];
        const uint32_t fb = net.fac_begin[r];
        const uint32_t fe = net.fac_begin[r + 1];
        for (uint32_t f = fb; f < fe; ++f) a = a * (x[net.fac_s[f]] - net.fac_c[f]);
        acc = acc + a;
        cum[r] = acc;
      }
    }
    tot = acc;
  };

  auto check_range = [&]() {
#pragma unroll
    for (int s = 0; s < MS; ++s)
      if (s < (int)S && !(__builtin_fabs(x[s]) <= REBOP_EXACT_LIMIT)) flags |= REBOP_FLAG_RANGE;
  };

  auto init = [&](uint64_t j) {
    jl = j;
    traj = run.traj_begin + j;
    if (run.st_x) {
#pragma unroll
      for (int s = 0; s < MS; ++s) x[s] = s < (int)S ? run.st_x[(uint64_t)s * n + j] : 0.0;
      t = run.st_t[j];
      draw = run.st_draw[j];
    } else {
      const int64_t *src = run.x0_per_traj ? run.x0 + j * S : run.x0;
#pragma unroll
      for (int s = 0; s < MS; ++s) x[s] = s < (int)S ? (double)src[s] : 0.0;
      t = 0.0;
      draw = 0;
    }
    check_range();
    draw_start = draw;
    draw_mark = draw;
    ov = 0;
    i = 0;
    T = grid_time(run, 0);
    props();
  };

  auto record = [&](uint64_t gi) {
    if (run.out) {
#pragma unroll
      for (int s = 0; s < MS; ++s)
        if (s < (int)S) run.out[(gi * S + s) * n + jl] = (int64_t)x[s];
    }
    check_range();
    if (draw - draw_mark >= REBOP_DRAWS_PER_CHECK) flags |= REBOP_FLAG_RANGE;
    draw_mark = draw;
  };

  auto finish = [&]() {
    if (run.final_x) {
#pragma unroll
      for (int s = 0; s < MS; ++s)
        if (s < (int)S) run.final_x[(uint64_t)s * n + jl] = (int64_t)x[s];
    }
    if (run.st_x) {
#pragma unroll
      for (int s = 0; s < MS; ++s)
        if (s < (int)S) run.st_x[(uint64_t)s * n + jl] = x[s];
      run.st_t[jl] = t;
      run.st_draw[jl] = draw;
    }
    ev_acc += (draw - draw_start) - ov;
    dr_acc += draw - draw_start;
    tr_acc += 1;
  };

  bool done = true;
  {
    const uint64_t j0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j0 < n) {
      init(j0);
      done = false;
    }
  }

  for (;;) {
    if (run.persistent) {
      // refill finished lanes: one atomic per wave
      const uint64_t need = __ballot(done);
      if (need != 0ull) {
        const uint32_t lane = __lane_id();
        const int leader = __ffsll((unsigned long long)need) - 1;
        unsigned long long base = 0ull;
        if ((int)lane == leader)
          base = atomicAdd(&run.counters[REBOP_CTR_QUEUE], (unsigned long long)__popcll(need));
        base = __shfl(base, leader);
        if (done) {
          const uint64_t rank = __popcll(need & ((1ull << lane) - 1ull));
          const uint64_t nj = total_threads + base + rank;
          if (nj < n) {
            init(nj);
            done = false;
          }
        }
      }
    }
    if (done) break;

    if (!(0.0 < tot)) {
      // absorbing state: the remaining samples all see the current counts
      for (; i <= nb; ++i) record(i);
      t = grid_time(run, nb);
      finish();
      done = true;
      continue;
    }
    const rebop_u32x4 blk = rebop_draw_block(run.seed, traj, draw);
    draw += 1;
    const double e = rebop_exp1_of(blk);
    const double tn = t + e / tot;
    if (tn > T) {
      t = T;
      record(i);
      ov += 1;
      i += 1;
      if (i > nb) {
        finish();
        done = true;
      } else {
        T = grid_time(run, i);
      }
      continue;
    }
    t = tn;
    const double chosen = tot * rebop_unif_of(blk);
    uint32_t k = 0;
#pragma unroll
    for (int r = 0; r < MR; ++r)
      if (r < (int)R) k += (cum[r] < chosen) ? 1u : 0u;
    const double *drow = lds_diff + (uint64_t)k * S;
#pragma unroll
    for (int s = 0; s < MS; ++s)
      if (s < (int)S) x[s] = x[s] + drow[s];
    props();
  }

  atomicAdd(&blk_ctr[0], (unsigned long long)ev_acc);
  atomicAdd(&blk_ctr[1], (unsigned long long)dr_acc);
  atomicAdd(&blk_ctr[2], (unsigned long long)tr_acc);
  if (flags) atomicOr(&blk_ctr[3], (unsigned long long)flags);
  __syncthreads();
  if (threadIdx.x == 0) {
    atomicAdd(&run.counters[REBOP_CTR_EVENTS], blk_ctr[0]);
    atomicAdd(&run.counters[REBOP_CTR_DRAWS], blk_ctr[1]);
    atomicAdd(&run.counters[REBOP_CTR_TRAJ], blk_ctr[2]);
    if (blk_ctr[3]) atomicOr(&run.counters[REBOP_CTR_FLAGS], blk_ctr[3]);
  }
}

// Clears the launch counters with a kernel (not hipMemsetAsync), so that the
// sequence stays valid inside a captured hipGraph.
__global__ void reset_counters_kernel(unsigned long long *counters) {
  if (threadIdx.x < REBOP_CTR_COUNT) counters[threadIdx.x] = 0ull;
}

} // namespace rebop

// ---------------------------------------------------------------------------
// launchers (C++ linkage, used by rebop_gpu.cpp)

namespace rebop {

struct RegShape {
  int ms, mr;
};

static const RegShape kRegShapes[] = {{4, 8}, {8, 16}, {16, 32}, {32, 64}};

int reg_shape_count() { return (int)(sizeof(kRegShapes) / sizeof(kRegShapes[0])); }
int reg_shape_ms(int c) { return kRegShapes[c].ms; }
int reg_shape_mr(int c) { return kRegShapes[c].mr; }

static const void *reg_kernel_ptr(int c) {
  switch (c) {
  case 0: return (const void *)&ssa_reg_kernel<4, 8>;
  case 1: return (const void *)&ssa_reg_kernel<8, 16>;
  case 2: return (const void *)&ssa_reg_kernel<16, 32>;
  case 3: return (const void *)&ssa_reg_kernel<32, 64>;
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

hipError_t launch_reg(int c, const rebop_dev_net &net, const rebop_dev_run &run, int blocks,
                      int threads, size_t lds, hipStream_t stream) {
  switch (c) {
  case 0: hipLaunchKernelGGL((ssa_reg_kernel<4, 8>), dim3(blocks), dim3(threads), lds, stream, net, run); break;
  case 1: hipLaunchKernelGGL((ssa_reg_kernel<8, 16>), dim3(blocks), dim3(threads), lds, stream, net, run); break;
  case 2: hipLaunchKernelGGL((ssa_reg_kernel<16, 32>), dim3(blocks), dim3(threads), lds, stream, net, run); break;
  case 3: hipLaunchKernelGGL((ssa_reg_kernel<32, 64>), dim3(blocks), dim3(threads), lds, stream, net, run); break;
  default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

} // namespace rebop
