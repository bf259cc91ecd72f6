// ssa_lane.h -- the per-lane SSA engine, shared by every kernel.
//
// One trajectory per lane.  Each lane runs rebop's grid loop
// (src/pyo3_gillespie.rs:200-208) around Gillespie::advance_until
// (src/gillespie.rs:315-344) as a state machine whose every iteration is one
// exponential draw (or, in an absorbing state, one sample without a draw):
//
//   tot <= 0 or NaN  -> absorbing: sample i sees the current counts, no draw
//                       (gillespie.rs:323-326)
//   t + E/tot > T_i  -> overshoot: t = T_i, sample i, the draw is discarded
//                       (gillespie.rs:327-331); counts are unchanged, so the
//                       cumulative propensities are reused for T_{i+1}
//   otherwise        -> event: chosen = tot*U, k = #{r : cum_r < chosen}
//                       (choose_cumrate_sum, gillespie.rs:402-407), apply
//                       reaction k (Jump::affect, :142-152); the propensities
//                       are recomputed in the reference's order (make_cumrates
//                       :357-364 over Rate::rate :71-90) at the next iteration
//
// The network enters through a policy class `Net` providing
//   static constexpr int MS, MR;          register capacity (species, reactions)
//   uint32_t S() const, R() const;
//   double props(const double (&x)[MS], double (&cum)[MR]) const;   -> total
//   uint32_t choose(const double (&cum)[MR], double chosen) const;
//   void affect(double (&x)[MS], uint32_t k) const;
// so that one loop serves both a data-driven network (any network within the
// capacity) and a network compiled into the kernel's code.
//
// Counts are exact FP64 integers in VGPRs (see ssa_layout.h for the range
// guard).  A persistent launch refills a finished lane with the next
// trajectory id from a device-wide counter: one atomic per wave per refill
// (ballot over the finished lanes, rank by popcount).
#ifndef REBOP_SSA_LANE_H
#define REBOP_SSA_LANE_H

#include "rebop_rng.h"
#include "ssa_layout.h"

#if defined(__clang__)
#pragma clang fp contract(off)
#endif

namespace rebop {

__device__ __forceinline__ double grid_time(const rebop_dev_run &run, uint64_t i) {
  if (run.mode == REBOP_MODE_TARGET) return run.tmax;
  // src/pyo3_gillespie.rs:202: tmax * i as f64 / nb_steps as f64
  return run.tmax * (double)i / (double)run.nb_steps;
}

struct LaneTotals {
  uint64_t events = 0, draws = 0;
  uint32_t trajectories = 0, flags = 0;
};

template <class Net>
__device__ __forceinline__ LaneTotals lane_loop(const Net &net, const rebop_dev_run &run) {
  constexpr int MS = Net::MS;
  constexpr int MR = Net::MR;
  const uint32_t S = net.S();
  const uint64_t n = run.n_traj;
  const uint32_t nb = run.mode == REBOP_MODE_TARGET ? 0u : (uint32_t)run.nb_steps;
  const uint64_t Sn = (uint64_t)S * n;
  const uint64_t gtid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t total_threads = (uint64_t)gridDim.x * blockDim.x;

  double x[MS];
  double cum[MR];
  double tot = 0.0, t = 0.0, T = 0.0;
  uint64_t draw = 0, draw0 = 0, mark = 0, jl = 0;
  int64_t *orec = nullptr;
  uint32_t i = 0, ov = 0;
  LaneTotals acc;
  bool done = true, started = false, need_props = false;

#pragma unroll
  for (int s = 0; s < MS; ++s) x[s] = 0.0;
#pragma unroll
  for (int r = 0; r < MR; ++r) cum[r] = 0.0;

  for (;;) {
    if (done) {
      // next trajectory: the lane's own id first, then (persistent launches)
      // ids from the device-wide queue
      uint64_t nj = n;
      if (!started) {
        nj = gtid;
        started = true;
      } else if (run.persistent) {
        const uint64_t mask = __ballot(1);
        const uint32_t lane = __lane_id();
        const int leader = __ffsll((unsigned long long)mask) - 1;
        unsigned long long base = 0ull;
        if ((int)lane == leader)
          base = atomicAdd(&run.counters[REBOP_CTR_QUEUE], (unsigned long long)__popcll(mask));
        base = __shfl(base, leader);
        nj = total_threads + base + (uint64_t)__popcll(mask & ((1ull << lane) - 1ull));
      }
      if (nj >= n) break;
      jl = nj;
      if (run.st_x) {
#pragma unroll
        for (int s = 0; s < MS; ++s) x[s] = s < (int)S ? run.st_x[(uint64_t)s * n + jl] : 0.0;
        t = run.st_t[jl];
        draw = run.st_draw[jl];
      } else {
        const int64_t *src = run.x0_per_traj ? run.x0 + jl * S : run.x0;
#pragma unroll
        for (int s = 0; s < MS; ++s) x[s] = s < (int)S ? (double)src[s] : 0.0;
        t = 0.0;
        draw = 0;
      }
#pragma unroll
      for (int s = 0; s < MS; ++s)
        if (s < (int)S && !(__builtin_fabs(x[s]) <= REBOP_EXACT_LIMIT)) acc.flags |= REBOP_FLAG_RANGE;
      draw0 = draw;
      mark = draw;
      ov = 0;
      i = 0;
      T = grid_time(run, 0);
      orec = run.out ? run.out + jl : nullptr;
      need_props = true;
      done = false;
    }

    if (need_props) {
      tot = net.props(x, cum);
      need_props = false;
    }

    bool rec = true; // absorbing state: sample without drawing
    double tn = 0.0, u = 0.0;
    if (0.0 < tot) {
      const rebop_u32x4 blk = rebop_draw_block(run.seed, run.traj_begin + jl, draw);
      draw += 1;
      tn = t + rebop_exp1_of(blk) / tot;
      u = rebop_unif_of(blk);
      rec = tn > T; // overshoot: the draw is discarded
      ov += rec ? 1u : 0u;
    }

    if (rec) {
      if (orec) {
#pragma unroll
        for (int s = 0; s < MS; ++s)
          if (s < (int)S) orec[(uint64_t)s * n] = (int64_t)x[s];
        orec += Sn;
      }
#pragma unroll
      for (int s = 0; s < MS; ++s)
        if (s < (int)S && !(__builtin_fabs(x[s]) <= REBOP_EXACT_LIMIT)) acc.flags |= REBOP_FLAG_RANGE;
      if (draw - mark >= REBOP_DRAWS_PER_CHECK) acc.flags |= REBOP_FLAG_RANGE;
      mark = draw;
      t = T;
      i += 1;
      if (i > nb) {
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
        acc.events += (draw - draw0) - ov;
        acc.draws += draw - draw0;
        acc.trajectories += 1;
        done = true;
      } else {
        T = grid_time(run, i);
      }
    } else {
      t = tn;
      const uint32_t k = net.choose(cum, tot * u);
      net.affect(x, k);
      need_props = true;
    }
  }
  return acc;
}

// Sums the lane totals of a workgroup in LDS, then once into the launch counters.
__device__ __forceinline__ void flush_totals(const LaneTotals &acc, const rebop_dev_run &run,
                                             unsigned long long *blk_ctr) {
  atomicAdd(&blk_ctr[0], (unsigned long long)acc.events);
  atomicAdd(&blk_ctr[1], (unsigned long long)acc.draws);
  atomicAdd(&blk_ctr[2], (unsigned long long)acc.trajectories);
  if (acc.flags) atomicOr(&blk_ctr[3], (unsigned long long)acc.flags);
  __syncthreads();
  if (threadIdx.x == 0) {
    atomicAdd(&run.counters[REBOP_CTR_EVENTS], blk_ctr[0]);
    atomicAdd(&run.counters[REBOP_CTR_DRAWS], blk_ctr[1]);
    atomicAdd(&run.counters[REBOP_CTR_TRAJ], blk_ctr[2]);
    if (blk_ctr[3]) atomicOr(&run.counters[REBOP_CTR_FLAGS], blk_ctr[3]);
  }
}

} // namespace rebop

#endif
