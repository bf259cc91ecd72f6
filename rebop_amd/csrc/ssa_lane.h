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
//                       :357-364 over Rate::rate :71-90) right after the jump
//
// A watchdog the reference does not have: with run.draw_cap set, a
// trajectory whose draws in this launch reach the cap is stopped at that
// event (its state is treated as absorbing from there), so a trajectory whose
// propensity keeps growing cannot hold its wave, and the launch, forever.
// The run then reports REBOP_FLAG_CAP.  The check is compiled in only for
// policies with Net::kCap (the host launches those when a cap is set), so
// uncapped runs pay nothing for it.
//
// The network enters through a policy object `net`.  The lane's register
// state is two local arrays of lane_loop, double x[Net::MS] (counts, for
// policies that keep them in VGPRs) and double cum[Net::MR] (what choose
// needs), handed to every policy call (local arrays, not struct members, so
// that the compiler keeps them in VGPRs even when x is indexed by a
// wave-uniform runtime species index):
//   static constexpr bool kCap;  compile the draw-cap checks in
//   static constexpr uint32_t kRangeFlag;  raised when in_range fails (REBOP_FLAG_EXACT
//                                          marks f64 counts, checked every 2^31 draws)
//   static constexpr uint32_t kStopFlag;   nonzero: once this flag is raised anywhere the
//                                          launch may be rerun (bool may_stop: it will),
//                                          so lanes stop early
//   uint32_t S() const;
//   void start(const rebop_dev_run &run, uint64_t j);    per-trajectory setup (rates)
//   void load_x0(X &x, const int64_t *src);                counts from an int64 row
//   void load_state(X &x, const double *st, uint64_t stride);  st[s * stride]
//   void store_state(const X &x, double *st, uint64_t stride);
//   void write_counts(const X &x, int64_t *dst, uint64_t stride);  dst[s * stride]
//   void write_raw(const X &x, int64_t *dst, uint64_t stride);     the samples: as
//        write_counts, or (counts in f64) their f64 bits, which a pass after
//        the launch converts in place (raw_counts_to_int64)
//   bool in_range(const X &x);          counts inside the exactly simulated range
//   double grid_time(const rebop_dev_run &run, uint64_t i);   sample time i
//   double props(X &x, C &cum);         make_cumrates -> total
//   uint32_t choose(X &x, C &cum, double chosen);   a reaction, or >= R for none
//   void affect(X &x, uint32_t k);                   (then no change)
// so that one loop serves counts in VGPRs (reg_* helpers below) and in LDS,
// for a data-driven network or one compiled into the kernel.
//
// A persistent launch refills a finished lane with the next trajectory id
// from a device-wide counter: one atomic per wave per refill (ballot over the
// finished lanes, rank by popcount).
#ifndef REBOP_SSA_LANE_H
#define REBOP_SSA_LANE_H

#include "rebop_rng.h"
#include "ssa_layout.h"

#if defined(__clang__)
#pragma clang fp contract(off)
#endif

namespace rebop {

__device__ __forceinline__ double grid_time(double tmax, uint64_t i, uint64_t nb) {
  // src/pyo3_gillespie.rs:202: tmax * i as f64 / nb_steps as f64
  return tmax * (double)i / (double)nb;
}

__device__ __forceinline__ double grid_time(const rebop_dev_run &run, uint64_t i) {
  if (run.mode == REBOP_MODE_TARGET) return run.tmax;
  return grid_time(run.tmax, i, run.nb_steps);
}

// The sample time from the table written before the launch (same value).
__device__ __forceinline__ double table_time(const rebop_dev_run &run, uint64_t i) {
  return run.grid[i];
}

struct LaneTotals {
  uint64_t events = 0, draws = 0;
  uint32_t trajectories = 0, flags = 0;
};

// Helpers for counts held in VGPRs as exact FP64 integers (species >= S stay 0).
template <int MS>
__device__ __forceinline__ void reg_load_x0(double (&x)[MS], const int64_t *src, uint32_t S) {
#pragma unroll
  for (int s = 0; s < MS; ++s) x[s] = s < (int)S ? (double)src[s] : 0.0;
}
template <int MS>
__device__ __forceinline__ void reg_load_state(double (&x)[MS], const double *st, uint64_t stride,
                                               uint32_t S) {
#pragma unroll
  for (int s = 0; s < MS; ++s) x[s] = s < (int)S ? st[(uint64_t)s * stride] : 0.0;
}
template <int MS>
__device__ __forceinline__ void reg_store_state(const double (&x)[MS], double *st, uint64_t stride,
                                                uint32_t S) {
#pragma unroll
  for (int s = 0; s < MS; ++s)
    if (s < (int)S) st[(uint64_t)s * stride] = x[s];
}
template <int MS>
__device__ __forceinline__ void reg_write_counts(const double (&x)[MS], int64_t *dst,
                                                 uint64_t stride, uint32_t S) {
#pragma unroll
  for (int s = 0; s < MS; ++s)
    if (s < (int)S) dst[(uint64_t)s * stride] = (int64_t)x[s];
}
// The sample's counts as their f64 bits (8 bytes each, in the int64 slots):
// the conversion, five FP64 operations per count, runs after the launch in
// a streaming pass instead of in the event loop.
template <int MS>
__device__ __forceinline__ void reg_write_raw(const double (&x)[MS], int64_t *dst, uint64_t stride,
                                              uint32_t S) {
#pragma unroll
  for (int s = 0; s < MS; ++s)
    if (s < (int)S) dst[(uint64_t)s * stride] = (int64_t)rebop_f64_bits(x[s]);
}

template <int MS>
__device__ __forceinline__ bool reg_in_range(const double (&x)[MS], uint32_t S) {
  bool ok = true;
#pragma unroll
  for (int s = 0; s < MS; ++s)
    if (s < (int)S) ok = ok && (__builtin_fabs(x[s]) <= REBOP_EXACT_LIMIT);
  return ok;
}

// The next local trajectory index of a lane: its own `first` the first time,
// then (persistent launches) queue_offset + an index from the device queue,
// one atomic per wave (ballot over the lanes asking, rank by popcount).
__device__ __forceinline__ uint64_t next_trajectory(const rebop_dev_run &run, bool &started,
                                                    uint64_t first, uint64_t queue_offset) {
  if (!started) {
    started = true;
    return first;
  }
  if (!run.persistent) return run.n_traj;
  const uint64_t mask = __ballot(1);
  const uint32_t lane = __lane_id();
  const int leader = __ffsll((unsigned long long)mask) - 1;
  unsigned long long base = 0ull;
  if ((int)lane == leader) base = atomicAdd(run.queue, (unsigned long long)__popcll(mask));
  base = __shfl(base, leader);
  return queue_offset + base + (uint64_t)__popcll(mask & ((1ull << lane) - 1ull));
}

// Whether some lane of this launch raised `flag` (a launch that will be
// rerun: the rest of its work would be thrown away).
__device__ __forceinline__ bool launch_stopped(const rebop_dev_run &run, uint32_t flag) {
  return (__hip_atomic_load(&run.counters[REBOP_CTR_FLAGS], __ATOMIC_RELAXED,
                            __HIP_MEMORY_SCOPE_AGENT) &
          flag) != 0ull;
}

// The draw count at which a trajectory that started this launch with d0
// draws hits the watchdog (never, without a cap).
__device__ __forceinline__ uint64_t cap_limit(const rebop_dev_run &run, uint64_t d0) {
  const uint64_t c = run.draw_cap;
  return (c == 0ull || d0 + c < d0) ? ~0ull : d0 + c;
}

// A trajectory reached the draw cap: flag the run, count it once (`lead`),
// and return the total propensity that makes its state absorbing from here.
__device__ __forceinline__ double hit_cap(const rebop_dev_run &run, LaneTotals &acc, bool lead) {
  acc.flags |= REBOP_FLAG_CAP;
  if (lead) atomicAdd(&run.counters[REBOP_CTR_CAPPED], 1ull);
  return 0.0;
}

// make_cumrates after a change of the counts.  An infinite total would stop
// time (E / tot == 0) for good: it flags the run and is returned as 0, so the
// trajectory runs out as an absorbing state and every lane finishes.
template <class Net, class X, class C>
__device__ __forceinline__ double props_checked(Net &net, X &x, C &cum, LaneTotals &acc) {
  double tot = net.props(x, cum);
  if (tot == __builtin_inf()) {
    acc.flags |= REBOP_FLAG_RANGE;
    tot = 0.0;
  }
  return tot;
}

// Runs trajectories on this lane until none is left.  `first` is the lane's
// first local trajectory index (>= run.n_traj for none); a persistent launch
// then takes indices queue_offset + atomicAdd(queue) from the device queue.
// `ln_tab`: the ln table, or the workgroup's copy of it in LDS.
template <class Net>
__device__ __forceinline__ LaneTotals lane_loop(Net &net, const rebop_dev_run &run, uint64_t first,
                                                uint64_t queue_offset,
                                                const double (*ln_tab)[4] = rebop_ln_table) {
  const uint64_t n = run.n_traj;
  const uint32_t nb = run.mode == REBOP_MODE_TARGET ? 0u : (uint32_t)run.nb_steps;
  const uint64_t Sn = (uint64_t)net.S() * n;

  double x[Net::MS];
  double cum[Net::MR];
#pragma unroll
  for (int s = 0; s < Net::MS; ++s) x[s] = 0.0;
#pragma unroll
  for (int r = 0; r < Net::MR; ++r) cum[r] = 0.0;
  double tot = 0.0, t = 0.0, T = 0.0;
  uint64_t draw = 0, draw0 = 0, mark = 0, jl = 0, dlim = ~0ull;
  int64_t *orec = nullptr;
  uint32_t i = 0, ov = 0;
  LaneTotals acc;
  bool done = true, started = false, spill = false;
  bool stop = false; // (kStopFlag) the launch will be rerun: take nothing more
  rebop_stream st = rebop_stream_of(run.seed, 0ull);

  for (;;) {
    bool leave = false;
    if (done) {
      // next trajectory: the lane's own first index, then (persistent
      // launches) indices from the device-wide queue
      if constexpr (Net::kStopFlag != 0u)
        stop = stop || (net.may_stop && launch_stopped(run, Net::kStopFlag));
      const uint64_t nj = stop ? n : next_trajectory(run, started, first, queue_offset);
      if (nj >= n) {
        leave = true;
      } else {
        jl = nj;
        st = rebop_stream_of(run.seed, run.traj_begin + jl);
        net.start(run, jl);
        if (run.st_x) {
          net.load_state(x, run.st_x + jl, n);
          t = run.st_t[jl];
          draw = run.st_draw[jl];
        } else {
          net.load_x0(x, run.x0_div ? run.x0 + (jl / run.x0_div) * net.S() : run.x0);
          t = 0.0;
          draw = 0;
        }
        if (!net.in_range(x)) acc.flags |= Net::kRangeFlag;
        draw0 = draw;
        mark = draw;
        if constexpr (Net::kCap) dlim = cap_limit(run, draw);
        ov = 0;
        i = 0;
        T = net.grid_time(run, 0);
        orec = run.out ? run.out + jl : nullptr;
        done = false;
        tot = props_checked(net, x, cum, acc);
      }
    }
    // Tail hand-off.  In a persistent launch a lane leaves the loop only when
    // the queue is empty, so a thin wave means no work is left to take: its
    // remaining lanes leave too and write their state out below, and a
    // grouped launch finishes those trajectories (lane_loop_group).  Both
    // reasons share one exit, which keeps the loop's scalar registers free.
    if (!leave && run.spill_below && (uint32_t)__popcll(__ballot(1)) <= run.spill_below) {
      spill = true;
      leave = true;
    }
    if (leave) break;

    bool rec = true; // absorbing state: sample without drawing
    double tn = 0.0, u = 0.0;
    if (0.0 < tot) {
      const rebop_u32x4 blk = rebop_stream_block(st, draw);
      draw += 1;
      tn = t + rebop_exp1_of_tab(blk, ln_tab) / tot;
      u = rebop_unif_of(blk);
      rec = tn > T; // overshoot: the draw is discarded
      ov += rec ? 1u : 0u;
    }

    if (rec) {
      if (orec) {
        net.write_raw(x, orec, n);
        orec += Sn;
      }
      if (!net.in_range(x)) acc.flags |= Net::kRangeFlag;
      if (Net::kRangeFlag == REBOP_FLAG_EXACT && draw - mark >= REBOP_DRAWS_PER_CHECK)
        acc.flags |= REBOP_FLAG_EXACT; // f64 counts: too many draws between two checks
      if constexpr (Net::kStopFlag != 0u) {
        // (a launch that will be rerun) this lane met the flag: tell the
        // others at once; else look every 8th sample whether another lane
        // did; either way drop the trajectory
        if (!net.may_stop) {
        } else if (acc.flags & Net::kStopFlag) {
          if (!stop) atomicOr(&run.counters[REBOP_CTR_FLAGS], (unsigned long long)Net::kStopFlag);
          stop = true;
        } else if ((i & 7u) == 7u) {
          stop = launch_stopped(run, Net::kStopFlag);
        }
        if (stop) {
          done = true;
          continue;
        }
      }
      mark = draw;
      t = T;
      i += 1;
      if (i > nb) {
        if (run.final_x) net.write_counts(x, run.final_x + jl, n);
        if (run.st_x) {
          net.store_state(x, run.st_x + jl, n);
          run.st_t[jl] = t;
          run.st_draw[jl] = draw;
        }
        acc.events += (draw - draw0) - ov;
        acc.draws += draw - draw0;
        acc.trajectories += 1;
        done = true;
      } else {
        T = net.grid_time(run, i);
      }
    } else {
      t = tn;
      net.affect(x, net.choose(x, cum, tot * u));
      tot = props_checked(net, x, cum, acc);
      if constexpr (Net::kCap)
        if (__builtin_expect(draw >= dlim, 0)) tot = hit_cap(run, acc, true);
    }
  }
  if (spill) {
    // one entry per handed-off trajectory, one atomic per wave
    const uint64_t act = __ballot(1);
    const uint32_t lane = __lane_id();
    const int leader = __ffsll((unsigned long long)act) - 1;
    unsigned long long base = 0ull;
    if ((int)lane == leader)
      base = atomicAdd(&run.counters[REBOP_CTR_SPILL], (unsigned long long)__popcll(act));
    base = __shfl(base, leader);
    const uint64_t e = base + (uint64_t)__popcll(act & ((1ull << lane) - 1ull));
    if (e < run.spill_cap) {
      run.spill_j[e] = jl;
      run.spill_draw[e] = draw;
      run.spill_mark[e] = mark;
      run.spill_t[e] = t;
      // bit 31: the trajectory was stopped by the draw cap (its zero total
      // is forced, not recomputed from the counts); grid positions < 2^31
      run.spill_i[e] = i | ((Net::kCap && tot == 0.0 && draw >= dlim) ? 0x80000000u : 0u);
      net.store_state(x, run.spill_x + e, run.spill_cap);
    } else {
      acc.flags |= REBOP_FLAG_INTERNAL; // cannot happen: one entry per lane
    }
    acc.events += (draw - draw0) - ov;
    acc.draws += draw - draw0;
  }
  return acc;
}

// Lane q of each aligned group of K lanes (K = 2 or 4) reads `v` from lane Q
// of its group: one DPP quad permutation per 32-bit half, no LDS traffic.
template <int K, int Q>
__device__ __forceinline__ double group_read(double v) {
  static_assert(K == 2 || K == 4, "groups of 2 or 4 lanes");
  constexpr int ctrl = K == 4 ? (Q | (Q << 2) | (Q << 4) | (Q << 6))
                              : (Q | (Q << 2) | ((2 + Q) << 4) | ((2 + Q) << 6));
  const uint64_t b = rebop_f64_bits(v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)b, ctrl, 0xF, 0xF, false);
  const uint32_t hi =
      (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)(b >> 32), ctrl, 0xF, 0xF, false);
  return rebop_f64_from_bits(((uint64_t)hi << 32) | lo);
}

template <int K>
__device__ __forceinline__ double group_read(double v, int q) {
  switch (q) {
  case 0: return group_read<K, 0>(v);
  case 1: return group_read<K, 1>(v);
  case 2: if constexpr (K == 4) return group_read<K, 2>(v); else return v;
  default: if constexpr (K == 4) return group_read<K, 3>(v); else return v;
  }
}

// K lanes per trajectory, for launches with too few trajectories to give
// every SIMD more than one wave (there the time is one trajectory's serial
// event chain, and the Philox block + ln of each draw is most of an event).
// Lane `sub` of a group computes the block of draw (draw + sub) of the group's
// trajectory; the group then runs up to K iterations of the state machine,
// iteration q using block q (read from lane q by DPP).  That is exactly the
// draw each iteration would take: every event or overshoot consumes the next
// draw, and a trajectory in an absorbing state never draws again (its counts
// no longer change), so the blocks of a round are used in order or not at
// all.  All lanes of a group hold the same state; lane 0 of the group writes
// and counts.  Plain launches only: trajectory j is group j, or with
// run.resume the trajectory of spill entry j (the tail of a persistent
// launch, continued from its written-out state).
//
// One wave per SIMD issues every instruction of the chain itself, so the
// loop keeps memory latency off that chain: `ln_tab` is the workgroup's LDS
// copy of the ln table, whose rows for the next round's draws are read while
// this round's events run (the next round's blocks are (draw + K + sub),
// since a round either consumes K draws or ends the drawing for good), and
// `grid_lds` its LDS copy of the sample times, read at every iteration while
// the division runs.  No vector-memory wait is left in the loop, so the
// sample stores are never waited for.  GridLds: grid_lds holds the sample
// times (else the next one is carried, computed at each sample, for grids too
// long for LDS).
template <class Net, int K, bool GridLds>
__device__ __forceinline__ LaneTotals lane_loop_group(Net &net, const rebop_dev_run &run,
                                                      uint64_t gtid, const double (*ln_tab)[4],
                                                      const double *grid_lds) {
  LaneTotals acc;
  const uint64_t n = run.n_traj;
  const uint64_t e = gtid / K; // group index
  uint64_t jl = e;
  if (run.resume) {
    const uint64_t cnt = __hip_atomic_load(&run.counters[REBOP_CTR_SPILL], __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
    if (e >= cnt || e >= run.spill_cap) return acc;
    jl = run.spill_j[e];
  } else if (jl >= n) {
    return acc;
  }
  const uint32_t sub = __lane_id() % K;
  const bool lead = sub == 0;
  const uint32_t nb = run.mode == REBOP_MODE_TARGET ? 0u : (uint32_t)run.nb_steps;
  const uint64_t Sn = (uint64_t)net.S() * n;

  double x[Net::MS];
  double cum[Net::MR];
#pragma unroll
  for (int s = 0; s < Net::MS; ++s) x[s] = 0.0;
#pragma unroll
  for (int r = 0; r < Net::MR; ++r) cum[r] = 0.0;
  net.start(run, jl);
  double t;
  uint64_t draw, mark;
  uint32_t i = 0, ov = 0;
  bool capped = false;
  if (run.resume) {
    net.load_state(x, run.spill_x + e, run.spill_cap);
    t = run.spill_t[e];
    draw = run.spill_draw[e];
    mark = run.spill_mark[e];
    i = run.spill_i[e] & 0x7fffffffu;
    capped = (run.spill_i[e] >> 31) != 0u;
  } else if (run.st_x) {
    net.load_state(x, run.st_x + jl, n);
    t = run.st_t[jl];
    draw = run.st_draw[jl];
    mark = draw;
  } else {
    net.load_x0(x, run.x0_div ? run.x0 + (jl / run.x0_div) * net.S() : run.x0);
    t = 0.0;
    draw = 0;
    mark = 0;
  }
  if (!net.in_range(x)) acc.flags |= Net::kRangeFlag;
  const uint64_t draw0 = draw;
  // the cap counts from the trajectory's draws at the start of this call
  // (a resumed tail entry started it in the persistent launch)
  const uint64_t dlim =
      cap_limit(run, run.resume ? (run.st_x ? run.st_draw[jl] : 0ull) : draw);
  double Tc = GridLds ? 0.0 : grid_time(run, i), tot = 0.0;
  int64_t *orec = run.out && lead ? run.out + jl + (uint64_t)i * Sn : nullptr;
  bool done = false;
  tot = props_checked(net, x, cum, acc);
  if (capped) tot = 0.0; // stopped by the cap before the hand-off (already counted)

  // this lane's draw of the next round: ln reduced and its table row read,
  // the rest of the ln evaluated when the round starts
  rebop_ln_pre pre;
  double u_next;
  auto fetch = [&](uint64_t d) {
    const rebop_u32x4 blk = rebop_draw_block(run.seed, run.traj_begin + jl, d + sub);
    pre = rebop_ln64_begin(rebop_u01_open_closed(blk.w[0], blk.w[1]), ln_tab);
    u_next = rebop_unif_of(blk);
  };
  fetch(draw);
  while (!done) {
    const double e_own = -rebop_ln64_finish(pre), u_own = u_next;
    const uint64_t d0 = draw;
#pragma unroll
    for (int q = 0; q < K; ++q) {
      const double e = group_read<K>(e_own, q);
      const double u = group_read<K>(u_own, q);
      if (!done) {
        const double T = GridLds ? grid_lds[i] : Tc; // i <= nb until done
        // an absorbing (or capped) state takes no draw and records; the
        // quotient is then unused (E / 0 is not trapped)
        const bool live = 0.0 < tot;
        const double tn = t + e / tot;
        const bool rec = !live || tn > T;
        draw += live ? 1u : 0u;
        ov += (live && rec) ? 1u : 0u;
        if (rec) {
          if (orec) {
            net.write_raw(x, orec, n);
            orec += Sn;
          }
          if (!net.in_range(x)) acc.flags |= Net::kRangeFlag;
          if (Net::kRangeFlag == REBOP_FLAG_EXACT && draw - mark >= REBOP_DRAWS_PER_CHECK)
            acc.flags |= REBOP_FLAG_EXACT;
          mark = draw;
          t = T;
          i += 1;
          if constexpr (!GridLds) Tc = grid_time(run, i);
          if (i > nb) {
            if (lead && run.final_x) net.write_counts(x, run.final_x + jl, n);
            if (lead && run.st_x) {
              net.store_state(x, run.st_x + jl, n);
              run.st_t[jl] = t;
              run.st_draw[jl] = draw;
            }
            if (lead) {
              acc.events += (draw - draw0) - ov;
              acc.draws += draw - draw0;
              acc.trajectories += 1;
            }
            done = true;
          }
        } else {
          t = tn;
          net.affect(x, net.choose(x, cum, tot * u));
          tot = props_checked(net, x, cum, acc);
          if constexpr (Net::kCap)
            if (__builtin_expect(draw >= dlim, 0)) tot = hit_cap(run, acc, lead);
        }
      }
      // after the first event's sample-time read, so that waiting for that
      // read does not wait for these
      if (q == 0) fetch(d0 + K);
    }
  }
  return acc;
}

// The nb_steps == 0 loop of Gillespie.run (src/pyo3_gillespie.rs:209-223):
// record (t, counts) before the first event and after every
// _advance_one_reaction (src/gillespie.rs:275-297) while t < tmax.  That step
// has no tmax check (the last event lies past tmax) and sets t = +inf in an
// absorbing state.  Without rec_offsets the lane only counts its records
// into n_records; with them it writes row r of trajectory j at
// rec_offsets[j] + r (the stream is deterministic, so both passes agree).
template <class Net>
__device__ __forceinline__ LaneTotals lane_events(Net &net, const rebop_dev_run &run,
                                                  uint64_t first, uint64_t queue_offset,
                                                  const double (*ln_tab)[4] = rebop_ln_table) {
  const uint64_t n = run.n_traj;
  const uint32_t S = net.S();
  double x[Net::MS];
  double cum[Net::MR];
#pragma unroll
  for (int s = 0; s < Net::MS; ++s) x[s] = 0.0;
#pragma unroll
  for (int r = 0; r < Net::MR; ++r) cum[r] = 0.0;
  LaneTotals acc;
  bool started = false;
  for (;;) {
    const uint64_t jl = next_trajectory(run, started, first, queue_offset);
    if (jl >= n) break;
    const rebop_stream st = rebop_stream_of(run.seed, run.traj_begin + jl);
    net.start(run, jl);
    net.load_x0(x, run.x0_div ? run.x0 + (jl / run.x0_div) * S : run.x0);
    if (!net.in_range(x)) acc.flags |= Net::kRangeFlag;
    const bool write = run.rec_offsets != nullptr;
    const uint64_t base = write ? run.rec_offsets[jl] : 0ull;
    // the caller's count bounds the writes: a count that does not match this
    // trajectory is reported (below), never written past
    const uint64_t cap = write ? run.n_records[jl] : 0ull;
    double t = 0.0;
    uint64_t draw = 0, rec = 0;
    const uint64_t dlim = cap_limit(run, 0ull);
    bool stop = false; // the draw cap was reached
    auto emit = [&]() {
      if (write && rec < cap) {
        run.rec_times[base + rec] = t;
        net.write_counts(x, run.rec_states + (base + rec) * S, 1);
      }
      rec += 1;
    };
    emit();
    while (t < run.tmax && !stop) {
      const double tot = net.props(x, cum);
      if (!(0.0 < tot)) {
        t = __builtin_inf();
      } else if (tot == __builtin_inf()) {
        acc.flags |= REBOP_FLAG_RANGE; // time would stop (E / tot == 0)
        t = __builtin_inf();
      } else {
        const rebop_u32x4 blk = rebop_stream_block(st, draw);
        draw += 1;
        t = t + rebop_exp1_of_tab(blk, ln_tab) / tot;
        net.affect(x, net.choose(x, cum, tot * rebop_unif_of(blk)));
        if (!net.in_range(x)) acc.flags |= Net::kRangeFlag;
        if constexpr (Net::kCap) {
          if (__builtin_expect(draw >= dlim, 0)) {
            hit_cap(run, acc, true);
            stop = true;
          }
        }
      }
      emit();
    }
    if (!write) run.n_records[jl] = rec;
    else if (run.n_records[jl] != rec) acc.flags |= REBOP_FLAG_RANGE;
    acc.events += draw;
    acc.draws += draw;
    acc.trajectories += 1;
  }
  return acc;
}

// lane_events with K lanes per trajectory (K = 2 or 4; plain launches:
// trajectory j is group j), for runs too small to give every SIMD more than
// one wave: lane `sub` of a group computes the Philox block and exponential
// of draw (draw + sub), and the group runs up to K events of the loop above
// using block q for its q-th event (read from lane q by DPP) -- each event
// takes the next draw and an absorbing state takes none, so the blocks of a
// round are used in order or not at all (the argument of lane_loop_group).
// All lanes hold the same state; lane 0 of the group writes and counts.
template <class Net, int K>
__device__ __forceinline__ LaneTotals lane_events_group(Net &net, const rebop_dev_run &run, uint64_t gtid,
                                                        const double (*ln_tab)[4]) {
  LaneTotals acc;
  const uint64_t jl = gtid / K;
  if (jl >= run.n_traj) return acc;
  const uint32_t sub = __lane_id() % K;
  const bool lead = sub == 0;
  const uint32_t S = net.S();
  double x[Net::MS];
  double cum[Net::MR];
#pragma unroll
  for (int s = 0; s < Net::MS; ++s) x[s] = 0.0;
#pragma unroll
  for (int r = 0; r < Net::MR; ++r) cum[r] = 0.0;
  net.start(run, jl);
  net.load_x0(x, run.x0_div ? run.x0 + (jl / run.x0_div) * S : run.x0);
  if (!net.in_range(x)) acc.flags |= Net::kRangeFlag;
  const bool write = run.rec_offsets != nullptr;
  const uint64_t base = write ? run.rec_offsets[jl] : 0ull;
  const uint64_t cap = write ? run.n_records[jl] : 0ull;
  double t = 0.0;
  uint64_t draw = 0, rec = 0;
  const uint64_t dlim = cap_limit(run, 0ull);
  bool stop = false; // the draw cap was reached
  auto emit = [&]() {
    if (write && lead && rec < cap) {
      run.rec_times[base + rec] = t;
      net.write_counts(x, run.rec_states + (base + rec) * S, 1);
    }
    rec += 1;
  };
  emit();
  while (t < run.tmax && !stop) {
    const rebop_u32x4 blk = rebop_draw_block(run.seed, run.traj_begin + jl, draw + sub);
    const double e_own = rebop_exp1_of_tab(blk, ln_tab), u_own = rebop_unif_of(blk);
#pragma unroll
    for (int q = 0; q < K; ++q) {
      const double e = group_read<K>(e_own, q);
      const double u = group_read<K>(u_own, q);
      if (t < run.tmax && !stop) {
        const double tot = net.props(x, cum);
        if (!(0.0 < tot)) {
          t = __builtin_inf();
        } else if (tot == __builtin_inf()) {
          acc.flags |= REBOP_FLAG_RANGE; // time would stop (E / tot == 0)
          t = __builtin_inf();
        } else {
          draw += 1;
          t = t + e / tot;
          net.affect(x, net.choose(x, cum, tot * u));
          if (!net.in_range(x)) acc.flags |= Net::kRangeFlag;
          if constexpr (Net::kCap) {
            if (__builtin_expect(draw >= dlim, 0)) {
              hit_cap(run, acc, lead);
              stop = true;
            }
          }
        }
        emit();
      }
    }
  }
  if (lead) {
    if (!write) run.n_records[jl] = rec;
    else if (run.n_records[jl] != rec) acc.flags |= REBOP_FLAG_RANGE;
    acc.events += draw;
    acc.draws += draw;
    acc.trajectories += 1;
  }
  return acc;
}

// Sums lane totals over the workgroup in LDS (blk_ctr zeroed before use).
__device__ __forceinline__ void add_totals(const LaneTotals &acc, unsigned long long *blk_ctr) {
  atomicAdd(&blk_ctr[0], (unsigned long long)acc.events);
  atomicAdd(&blk_ctr[1], (unsigned long long)acc.draws);
  atomicAdd(&blk_ctr[2], (unsigned long long)acc.trajectories);
  if (acc.flags) atomicOr(&blk_ctr[3], (unsigned long long)acc.flags);
}

// After a workgroup barrier: one global update per workgroup.
__device__ __forceinline__ void flush_block_totals(unsigned long long *counters,
                                                   const unsigned long long *blk_ctr) {
  if (threadIdx.x == 0) {
    atomicAdd(&counters[REBOP_CTR_EVENTS], blk_ctr[0]);
    atomicAdd(&counters[REBOP_CTR_DRAWS], blk_ctr[1]);
    atomicAdd(&counters[REBOP_CTR_TRAJ], blk_ctr[2]);
    if (blk_ctr[3]) atomicOr(&counters[REBOP_CTR_FLAGS], blk_ctr[3]);
  }
}

// One global update per wave, reduced across the wave's lanes (no LDS): for
// kernels whose LDS is all spoken for.
__device__ __forceinline__ void flush_wave_totals(const LaneTotals &acc,
                                                  unsigned long long *counters) {
  unsigned long long ev = acc.events, dr = acc.draws, tr = acc.trajectories;
  uint32_t fl = acc.flags;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    ev += __shfl_xor(ev, off);
    dr += __shfl_xor(dr, off);
    tr += __shfl_xor(tr, off);
    fl |= __shfl_xor(fl, off);
  }
  if (__lane_id() == 0) {
    atomicAdd(&counters[REBOP_CTR_EVENTS], ev);
    atomicAdd(&counters[REBOP_CTR_DRAWS], dr);
    atomicAdd(&counters[REBOP_CTR_TRAJ], tr);
    if (fl) atomicOr(&counters[REBOP_CTR_FLAGS], (unsigned long long)fl);
  }
}

// The whole-launch epilogue of a single-network kernel.
__device__ __forceinline__ void flush_totals(const LaneTotals &acc, const rebop_dev_run &run,
                                             unsigned long long *blk_ctr) {
  add_totals(acc, blk_ctr);
  __syncthreads();
  flush_block_totals(run.counters, blk_ctr);
}

} // namespace rebop

#endif
