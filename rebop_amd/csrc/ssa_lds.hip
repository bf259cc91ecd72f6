// ssa_lds.hip -- LDS-resident SSA kernel for large and heterogeneous networks.
//
// Counts live in LDS, one column per lane (x[s] at lds[s * 64 + lane]), so
// every access -- uniform or lane-varying species index -- maps lane l to
// the same banks whatever the species: conflict-free (int16 counts: lanes 2i
// and 2i+1 share a bank, so lane-varying species cost at most a 2-way
// conflict).  The compact kernel keeps them as int16 (|x| <= 2^15 - 1, a
// quarter of the LDS of int64), the narrow one as int32 (|x| <= 2^30); LDS
// per wave is what bounds the waves per CU, so the narrower counts run more
// waves.  The wide one keeps int64, the reference's isize, for populations
// beyond int32 and beyond the f64 counts of the register and specialised
// kernels (|x| <= 2^52), with every addition checked for overflow.  The
// network (a "job") is read at wave-uniform addresses through the scalar
// cache, except in the per-lane block scan below.
//
// make_cumrates (src/gillespie.rs:357-364) runs as one uniform pass that
// keeps the running sum at the end of every block of B reactions (cum[b],
// B = ceil(R / MB)).  choose_cumrate_sum (:402-407) counts the reactions with
// cum_r < chosen; while every propensity is >= 0 the cumulative sum is
// non-decreasing and that count is the first r with cum_r >= chosen, so a
// lane finds its block from the checkpoints and re-adds only that block's
// propensities, starting from the checkpoint before it: the same sequence of
// additions as the pass, hence the same values.  A lane whose counts can be
// negative (negative initial counts, a network whose jump consumes more than
// its rate's order, a negative rate constant, or a count that went negative)
// uses the full linear count instead, which is the reference's definition.
//
// The network is read from LDS too: when a wave starts an item it copies the
// job's reaction records (rate constant and the first two propensity factors,
// 16 bytes each, see ssa_layout.h) next to the counts; the jump records (up
// to four entries, 16 bytes) stay in global memory, read once per event.  A make_cumrates pass then costs one uniform 16-byte LDS read and
// two count reads per reaction, independent from one reaction to the next
// (the loop is unrolled so the reads of later reactions overlap the
// arithmetic of earlier ones); a missing factor multiplies by exactly 1.0.
// Reactions whose factors or jump do not fit a record use the global lists.
//
// Work items are tickets of a job (ceil(n_traj / 64) per job); a wave takes
// items from a device-wide queue, so one launch can run thousands of
// different networks (heterogeneous batch) with the items ordered
// largest-first by the host.  A wave holding a ticket takes 64 trajectory
// indices from its job's own queue, and each lane whose trajectory ends takes
// the job's next index (one atomic per wave per refill), until the job's
// queue is empty: a lane waits for the wave's slowest trajectory only at the
// end of a job, not at the end of every group of 64.  Every trajectory is
// claimed exactly once, so the results do not depend on which wave runs it.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "rebop_expr.h"
#include "rebop_rng.h"
#include "ssa_lane.h"
#include "ssa_layout.h"

#pragma clang fp contract(off)

namespace rebop {

#if defined(__HIP_DEVICE_COMPILE__)
// tables read in place as global memory: vector loads counted on the vector
// memory counter only (a flat load also counts on the LDS one, so waiting
// for a group's counts in LDS would wait for the next group's records too)
typedef const __attribute__((address_space(1))) uint4 *rebop_grec_t;
#endif

template <int MB, int W, bool EXPR, bool GREC>
struct LdsNet {
  // counts of W bytes: int16 (compact), int32 (narrow) or the reference's
  // full int64 (wide)
  static constexpr bool WIDE = W == 8;
  using CountT = typename std::conditional<
      W == 8, int64_t, typename std::conditional<W == 4, int32_t, int16_t>::type>::type;
  static constexpr int MS = 1;  // counts are in LDS, not in x[]
  static constexpr bool kCap = true; // the check costs nothing next to a pass over R reactions
  // a compact or narrow kernel's range is the smaller one: the host reruns
  // such a run with wider counts; the wide one flags an int64 overflow, or a
  // count beyond the f64 state arrays' exact range in a resumable run
  static constexpr uint32_t kRangeFlag = WIDE ? REBOP_FLAG_RANGE : REBOP_FLAG_NARROW;
  static constexpr uint32_t kStopFlag = WIDE ? 0u : REBOP_FLAG_NARROW;
  static constexpr int64_t kLimit = W == 8   ? (int64_t)REBOP_EXACT_LIMIT
                                    : W == 4 ? (int64_t)REBOP_LDS_LIMIT
                                             : (int64_t)REBOP_LDS16_LIMIT;
  static constexpr int MR = MB; // cum[] holds the block checkpoints
  // job fields (wave-uniform)
  uint32_t nS, R, B, nblk;
  bool job_linear;
  bool job_slow; // some reaction needs the global factor lists
  const double *k;
  const uint32_t *fb;
  const uint32_t *fs;
  const double *fc;
  const uint32_t *jb;
  const uint32_t *js;
  const int32_t *jd;
  // Rate::Expr programs (EXPR kernels only)
  const uint32_t *eb;
  const uint32_t *ecode;
  const double *econst;
  // define_system reactant terms (NULL: Gillespie semantics) and its choice rule
  const uint32_t *sb;
  const uint32_t *ssp;
  const uint32_t *smul;
  bool job_system;
  // this lane's column of counts: xr at the row of ones, xl at species 0
  static constexpr uint32_t kBL = 64; // lanes per workgroup (one wave)
  CountT *xr;
  CountT *xl;
  // the job's reaction records in LDS, jump records in global memory
  const uint32_t *rec;
  const uint32_t *jrec;
  // the run keeps its state in f64 arrays (resumable batch): counts must stay
  // within |x| <= 2^52 there
  bool state_f64;
  // the host reruns this launch wider if a count leaves its range: lanes
  // may then drop their work (kStopFlag)
  bool may_stop;
  // per-lane flags
  bool linear;
  bool bad;

  __device__ __forceinline__ CountT &cnt(uint32_t s) const { return xl[s * kBL]; }
  // the count a record's factor word names: row 0 is a row of ones (an
  // absent factor), row s + 1 is species s
  __device__ __forceinline__ CountT rcnt(uint32_t w) const { return xr[(w >> 16) * kBL]; }
  __device__ __forceinline__ double fcnt(uint32_t s) const { return (double)cnt(s); }

  __device__ __forceinline__ uint32_t S() const { return nS; }
  __device__ __forceinline__ void start(const rebop_dev_run &, uint64_t) {}
  // each job has its own horizon: computed, not tabled
  __device__ __forceinline__ double grid_time(const rebop_dev_run &run, uint64_t i) const {
    return rebop::grid_time(run, i);
  }

  // whether a count is inside this kernel's range
  __device__ __forceinline__ bool fits(int64_t v) const {
    if constexpr (WIDE) return !state_f64 || (v <= kLimit && v >= -kLimit);
    else return v <= kLimit && v >= -kLimit;
  }
  __device__ __forceinline__ void set_count(uint32_t s, int64_t v) {
    if (!fits(v)) {
      bad = true;
      v = 0;
    }
    cnt(s) = (CountT)v;
    if (v < 0) linear = true;
  }
  __device__ __forceinline__ void load_x0(double (&)[MS], const int64_t *src) {
    linear = job_linear;
    bad = false;
    for (uint32_t s = 0; s < nS; ++s) set_count(s, src[s]);
  }
  // a resumable state: exact integers in f64 (|x| <= 2^52, as stored)
  __device__ __forceinline__ void load_state(double (&)[MS], const double *st, uint64_t stride) {
    linear = job_linear;
    bad = false;
    for (uint32_t s = 0; s < nS; ++s) {
      const double v = st[(uint64_t)s * stride];
      if (__builtin_fabs(v) <= REBOP_EXACT_LIMIT) {
        set_count(s, (int64_t)v);
      } else {
        bad = true;
        cnt(s) = 0;
      }
    }
  }
  __device__ __forceinline__ void store_state(const double (&)[MS], double *st, uint64_t stride) const {
    for (uint32_t s = 0; s < nS; ++s) st[(uint64_t)s * stride] = fcnt(s);
  }
  // integer counts: written as int64 directly (no conversion pass)
  __device__ __forceinline__ void write_raw(const double (&x)[MS], int64_t *dst, uint64_t stride) const {
    write_counts(x, dst, stride);
  }
  __device__ __forceinline__ void write_counts(const double (&)[MS], int64_t *dst, uint64_t stride) const {
    for (uint32_t s = 0; s < nS; ++s) dst[(uint64_t)s * stride] = (int64_t)cnt(s);
  }
  __device__ __forceinline__ bool in_range(const double (&)[MS]) const { return !bad; }

  // Expr::eval (src/expr.rs:24-38) of a postfix program (rebop_expr.h):
  // counts enter as f64, the operations in the program's order.  The host
  // checked that the program never underflows its stack nor exceeds
  // REBOP_EXPR_MAX_DEPTH values and leaves exactly one.
  __device__ double expr_rate(uint32_t b, uint32_t e) const {
    double st[REBOP_EXPR_MAX_DEPTH];
    uint32_t sp = 0;
    for (uint32_t i = b; i < e; ++i) {
      const uint32_t tok = ecode[i];
      const uint32_t op = REBOP_EXPR_OPCODE(tok), arg = REBOP_EXPR_ARG(tok);
      if (op == REBOP_OP_CONST) {
        st[sp++] = econst[arg];
      } else if (op == REBOP_OP_SPECIES) {
        st[sp++] = fcnt(arg);
      } else if (op == REBOP_OP_NEG) {
        st[sp - 1] = -st[sp - 1];
      } else if (op == REBOP_OP_EXP) {
        st[sp - 1] = rebop_exp64(st[sp - 1]);
      } else {
        const double y = st[--sp], x = st[sp - 1];
        double v;
        switch (op) {
        case REBOP_OP_ADD: v = x + y; break;
        case REBOP_OP_SUB: v = x - y; break;
        case REBOP_OP_MUL: v = x * y; break;
        case REBOP_OP_DIV: v = x / y; break;
        case REBOP_OP_POW: v = rebop_pow64(x, y); break;
        case REBOP_OP_MAX: v = rebop_fmax64(x, y); break;
        default: v = rebop_fmin64(x, y); break;
        }
        st[sp - 1] = v;
      }
    }
    return st[0];
  }

  // _rate_lma! of a define_system term (src/gillespie_macro.rs:128-142):
  // x (x-1) ... (x-m+1) as a wrapping isize product, converted once; the
  // multiplicity 0 of `0 A` still gives x (the macro's loop 1..0 is empty)
  __device__ __forceinline__ double system_lma(uint32_t s, uint32_t m) const {
    const int64_t x = (int64_t)cnt(s);
    uint64_t p = (uint64_t)x;
    for (uint32_t i = 1; i < m; ++i) p = p * (uint64_t)(x - (int64_t)i);
    return (double)(int64_t)p;
  }

  // Rate::rate of reaction r from the global factor lists (src/gillespie.rs:71-90),
  // its rate expression, or its define_system product (gillespie_macro.rs:108:
  // the rate times the terms' products, left to right)
  __device__ __forceinline__ double rate_slow(uint32_t r) const {
    if constexpr (EXPR) {
      const uint32_t b = eb[r], e = eb[r + 1];
      if (b != e) return expr_rate(b, e);
    }
    if (sb) {
      double a = k[r];
      const uint32_t te = sb[r + 1];
      for (uint32_t i = sb[r]; i < te; ++i) a = a * system_lma(ssp[i], smul[i]);
      return a;
    }
    double a = k[r];
    const uint32_t fe = fb[r + 1];
    for (uint32_t f = fb[r]; f < fe; ++f) a = a * (double)((int64_t)cnt(fs[f]) - (int64_t)fc[f]);
    return a;
  }

  // factor (x[s] - c) of a record word whose count x is read: the
  // reference's integer difference (int32 in the compact and narrow kernels,
  // |x| <= 2^30 and c < 4096; int64 in the wide one), converted once.  An
  // absent factor names the row of ones with c = 0: exactly 1.0, with no
  // test
  __device__ __forceinline__ double factor_val(uint32_t w, CountT x) const {
    if constexpr (WIDE) {
      return (double)(x - (int64_t)(w & 0xfffu));
    } else {
      return (double)((int32_t)x - (int32_t)(w & 0xfffu));
    }
  }
  __device__ __forceinline__ double factor(uint32_t w) const { return factor_val(w, rcnt(w)); }
  // record r (read in place: as global memory)
  __device__ __forceinline__ uint4 record(uint32_t r) const {
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (GREC) return ((rebop_grec_t)(const void *)rec)[r];
#endif
    return *(const uint4 *)(rec + (size_t)r * REBOP_REC_WORDS);
  }

  // Rate::rate of reaction r, factors in stored order: k * f1 * f2
  __device__ __forceinline__ double rate(uint32_t r) const {
    const uint4 q = record(r);
    if (q.w & REBOP_REC_SLOW_FACTORS) return rate_slow(r);
    double a = rebop_f64_from_bits(((uint64_t)q.y << 32) | q.x);
    a = a * factor(q.z);
    a = a * factor(q.w);
    return a;
  }

  // Rate::rate of reaction r when its factors are all in the record
  __device__ __forceinline__ double rate_fast(uint32_t r) const {
    const uint4 q = record(r);
    double a = rebop_f64_from_bits(((uint64_t)q.y << 32) | q.x);
    a = a * factor(q.z);
    a = a * factor(q.w);
    return a;
  }

  // k * f1 * f2 of one record whose factor counts are already read
  __device__ __forceinline__ double rate_of(const uint4 &q, CountT x1, CountT x2) const {
    double a = rebop_f64_from_bits(((uint64_t)q.y << 32) | q.x);
    a = a * factor_val(q.z, x1);
    a = a * factor_val(q.w, x2);
    return a;
  }

  // make_cumrates with a checkpoint at the end of every block
  __device__ __forceinline__ double props(double (&)[MS], double (&cum)[MR]) const {
    double acc = 0.0;
    uint32_t r = 0;
    if (R == 0) return 0.0; // no record to prefetch
    if (!job_slow) {
      // every reaction's factors are in its record: groups of four reactions
      // (blocks hold a multiple of 4), software-pipelined: the records of the
      // next group are read while this group's counts are in flight, so a
      // group costs one LDS round trip (the sum still adds them one by one)
      // the record tables end with four records of padding, so the
      // prefetch of the next group may read up to R + 3 (what it reads past
      // R - 1 is never used), and the four reads share one base address
      uint4 q0 = record(0), q1 = record(1u), q2 = record(2u), q3 = record(3u);
#pragma unroll
      for (int b = 0; b < MB; ++b) { // static: cum[] stays in VGPRs
        if ((uint32_t)b < nblk) {
          const uint32_t re = (r + B < R) ? r + B : R;
          for (; r + 4 <= re; r += 4) {
            const uint4 c0 = q0, c1 = q1, c2 = q2, c3 = q3;
            q0 = record(r + 4u);
            q1 = record(r + 5u);
            q2 = record(r + 6u);
            q3 = record(r + 7u);
            const CountT x0a = rcnt(c0.z), x0b = rcnt(c0.w);
            const CountT x1a = rcnt(c1.z), x1b = rcnt(c1.w);
            const CountT x2a = rcnt(c2.z), x2b = rcnt(c2.w);
            const CountT x3a = rcnt(c3.z), x3b = rcnt(c3.w);
            // keep the next records' reads ahead of the wait for these counts
            __builtin_amdgcn_sched_barrier(0);
            const double a0 = rate_of(c0, x0a, x0b), a1 = rate_of(c1, x1a, x1b);
            const double a2 = rate_of(c2, x2a, x2b), a3 = rate_of(c3, x3a, x3b);
            acc = acc + a0;
            acc = acc + a1;
            acc = acc + a2;
            acc = acc + a3;
          }
          for (; r < re; ++r) acc = acc + rate_fast(r);
          cum[b] = acc;
        }
      }
      return acc;
    }
#pragma unroll
    for (int b = 0; b < MB; ++b) {
      if ((uint32_t)b < nblk) {
        const uint32_t re = (r + B < R) ? r + B : R;
        for (; r < re; ++r) acc = acc + rate(r);
        cum[b] = acc;
      }
    }
    return acc;
  }

  __device__ __forceinline__ uint32_t choose(double (&)[MS], double (&cum)[MR], double chosen) const {
    if (job_system) return choose_first_above(cum, chosen);
    if (linear) {
      // choose_cumrate_sum as defined: count every r with cum_r < chosen
      double acc = 0.0;
      uint32_t kk = 0;
      for (uint32_t r = 0; r < R; ++r) {
        acc = acc + rate(r);
        kk += (acc < chosen) ? 1u : 0u;
      }
      return kk;
    }
    // non-decreasing sums: the block holding the first cum_r >= chosen
    uint32_t b = 0;
    double start = 0.0;
#pragma unroll
    for (int q = 0; q < MB; ++q) {
      if ((uint32_t)q < nblk && cum[q] < chosen) {
        b = (uint32_t)q + 1u;
        start = cum[q];
      }
    }
    if (b >= nblk) b = nblk - 1u; // unreachable: cum[nblk-1] = total > chosen
    uint32_t r = b * B;
    const uint32_t re = (r + B < R) ? r + B : R;
    double acc = start;
    for (; r < re; ++r) {
      acc = acc + rate(r);
      if (!(acc < chosen)) break;
    }
    return r < re ? r : re - 1u;
  }

  // _choice! (src/gillespie_macro.rs:146-170): the first r with chosen <
  // cum_r, strict, and R (no reaction) when there is none
  __device__ uint32_t choose_first_above(const double (&cum)[MR], double chosen) const {
    if (linear) {
      double acc = 0.0;
      for (uint32_t r = 0; r < R; ++r) {
        acc = acc + rate(r);
        if (chosen < acc) return r;
      }
      return R;
    }
    // non-decreasing sums: the first block whose end sum exceeds chosen
    uint32_t b = nblk;
    double start = 0.0;
#pragma unroll
    for (int q = MB - 1; q >= 0; --q) {
      if ((uint32_t)q < nblk && chosen < cum[q]) {
        b = (uint32_t)q;
        start = q > 0 ? cum[q - 1] : 0.0;
      }
    }
    if (b >= nblk) return R;
    uint32_t r = b * B;
    const uint32_t re = (r + B < R) ? r + B : R;
    double acc = start;
    for (; r < re; ++r) {
      acc = acc + rate(r);
      if (chosen < acc) return r;
    }
    return R; // unreachable: the block's end sum exceeds chosen
  }

  __device__ __forceinline__ void add_count(uint32_t s, int32_t d) {
    if constexpr (WIDE) {
      int64_t v;
      if (__builtin_add_overflow(cnt(s), (int64_t)d, &v)) bad = true; // the isize would overflow
      cnt(s) = v;
      if (v < 0) linear = true;
      if (!fits(v)) bad = true;
    } else {
      // |x| <= 2^30 and |d| < 2^23: no int32 wrap before the check
      const int32_t v = (int32_t)((uint32_t)(int32_t)cnt(s) + (uint32_t)d);
      cnt(s) = (CountT)v;
      if (v < 0) linear = true;
      if (v > (int32_t)kLimit || v < -(int32_t)kLimit) bad = true;
    }
  }

  // Jump::affect (src/gillespie.rs:142-152), sparse form; kk >= R (a
  // define_system choice that matched no reaction) changes nothing
  __device__ __forceinline__ void affect(double (&)[MS], uint32_t kk) {
    if (kk >= R) return;
#if defined(__HIP_DEVICE_COMPILE__)
    const uint4 j = ((rebop_grec_t)(const void *)jrec)[kk];
#else
    const uint4 j = *(const uint4 *)(jrec + (size_t)kk * REBOP_JREC_WORDS);
#endif
    if (record(kk).w & REBOP_REC_SLOW_JUMP) {
      const uint32_t e1 = jb[kk + 1];
      for (uint32_t e = jb[kk]; e < e1; ++e) add_count(js[e], jd[e]);
      return;
    }
    const uint32_t e[4] = {j.x, j.y, j.z, j.w};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int32_t d = (int32_t)e[q] >> 8; // arithmetic shift: the signed 24-bit difference
      if (d != 0) add_count(e[q] & 0xffu, d);
    }
  }
};

// At most 256 VGPRs, so that two waves share a SIMD: left to itself the
// compiler takes a few more (the checkpoints alone hold 64) and one wave per
// SIMD remains.  The int16 and int32 kernels fit without spilling.
//
// GREC: the reaction records are read where they lie in global memory
// (through the caches) instead of from an LDS copy: that leaves more LDS
// per CU for counts, which pays where the LDS, not the registers, bounds the
// waves per CU (the host decides).
template <int MB, int W, bool EXPR, bool GREC>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) ssa_lds_kernel(const rebop_dev_job *__restrict__ jobs,
                                                      const rebop_dev_item *__restrict__ items,
                                                      uint32_t n_items,
                                                      unsigned long long *counters,
                                                      unsigned long long *job_events,
                                                      unsigned long long *job_queue,
                                                      uint32_t rec_word_off, uint32_t rerun) {
  // no static LDS: the dynamic size (counts + records) alone decides how
  // many workgroups share a CU (e.g. 4 x 40 KB for 128 species, 512 reactions)
  extern __shared__ int32_t lds_x[]; // counts [1 + S_max][blockDim.x], then the records

  const uint32_t lane = __lane_id();
  using Net = LdsNet<MB, W, EXPR, GREC>;
  // row 0: ones, the count an absent factor reads (never written again)
  reinterpret_cast<typename Net::CountT *>(lds_x)[threadIdx.x] = (typename Net::CountT)1;
  LaneTotals total;
  for (;;) {
    // a compact or narrow launch that met a count beyond its range is rerun
    // wider by the host (`rerun`: its counters are read): take no further
    // item (the rerun rewrites everything)
    if (W != 8 && rerun &&
        (__hip_atomic_load(&counters[REBOP_CTR_FLAGS], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) &
         REBOP_FLAG_NARROW))
      break;
    // one item per wave; readfirstlane makes the index (and so every job
    // field and network read below) wave-uniform: scalar loads into SGPRs
    unsigned long long got = 0ull;
    if (lane == 0) got = atomicAdd(&counters[REBOP_CTR_QUEUE], 1ull);
    const uint32_t it = __builtin_amdgcn_readfirstlane(
        (uint32_t)(got < (unsigned long long)n_items ? got : (unsigned long long)n_items));
    if (it >= n_items) break;
    const rebop_dev_item item = items[it];
    const rebop_dev_job &job = jobs[item.job];
    // the wave's first 64 trajectories of the job; none left: next ticket
    unsigned long long base = 0ull;
    if (lane == 0) base = atomicAdd(&job_queue[item.job], 64ull);
    base = __shfl(base, 0);
    if (base >= job.n_traj) continue;
    // the job's reaction records into LDS (every item: a wave's consecutive
    // items may belong to different jobs), or read in place
    const uint32_t *lrec = job.rec;
    if constexpr (!GREC) {
      uint32_t *l = (uint32_t *)lds_x + rec_word_off;
      __syncthreads(); // the previous item's reads are done
      for (uint32_t w = threadIdx.x * 4u; w < job.R * REBOP_REC_WORDS; w += blockDim.x * 4u)
        *(uint4 *)(l + w) = *(const uint4 *)(job.rec + w);
      __syncthreads();
      lrec = l;
    }

    LdsNet<MB, W, EXPR, GREC> net;
    net.nS = job.S;
    net.R = job.R;
    net.B = job.block;
    net.nblk = job.nblocks;
    net.job_linear = (job.flags & REBOP_JOB_LINEAR) != 0u;
    net.job_slow = (job.flags & REBOP_JOB_SLOW) != 0u;
    net.k = job.k;
    net.fb = job.fac_begin;
    net.fs = job.fac_s;
    net.fc = job.fac_c;
    net.jb = job.jump_begin;
    net.js = job.jump_s;
    net.jd = job.jump_d;
    net.eb = job.expr_begin;
    net.ecode = job.expr_code;
    net.econst = job.expr_const;
    net.sb = job.sys_begin;
    net.ssp = job.sys_species;
    net.smul = job.sys_mult;
    net.job_system = (job.flags & REBOP_JOB_SYSTEM) != 0u;
    net.state_f64 = job.st_x != nullptr;
    net.xr = reinterpret_cast<typename Net::CountT *>(lds_x) + threadIdx.x;
    net.xl = net.xr + Net::kBL;
    net.rec = lrec;
    net.jrec = job.jrec;
    net.linear = net.job_linear;
    net.bad = false;
    net.may_stop = rerun != 0u;

    rebop_dev_run run{};
    run.seed = job.seed;
    run.traj_begin = job.traj_begin;
    run.n_traj = job.n_traj;
    run.tmax = job.tmax;
    run.nb_steps = job.nb_steps;
    run.mode = job.mode;
    run.persistent = 1; // lanes refill from the job's queue
    run.draw_cap = job.draw_cap;
    run.x0 = job.x0;
    run.x0_div = job.x0_div;
    run.out = job.out;
    run.final_x = job.final_x;
    run.st_x = job.st_x;
    run.st_t = job.st_t;
    run.st_draw = job.st_draw;
    run.counters = counters;
    run.queue = &job_queue[item.job];
    run.n_records = job.n_records;
    run.rec_offsets = job.rec_offsets;
    run.rec_times = job.rec_times;
    run.rec_states = job.rec_states;

    const uint64_t first = base + lane < job.n_traj ? base + lane : job.n_traj;
    const LaneTotals acc = run.mode == REBOP_MODE_EVENTS ? lane_events(net, run, first, 0ull)
                                                         : lane_loop(net, run, first, 0ull);
    if (job_events) {
      // events of this item, summed over the wave, one atomic
      unsigned long long e = acc.events;
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) e += __shfl_xor(e, off);
      if (lane == 0) atomicAdd(&job_events[item.job], e);
    }
    total.events += acc.events;
    total.draws += acc.draws;
    total.trajectories += acc.trajectories;
    total.flags |= acc.flags;
    if (W != 8 && rerun && __ballot((acc.flags & REBOP_FLAG_NARROW) != 0u) != 0ull && lane == 0)
      atomicOr(&counters[REBOP_CTR_FLAGS], (unsigned long long)REBOP_FLAG_NARROW); // stop the others
  }
  flush_wave_totals(total, counters);
}

// ---------------------------------------------------------------------------
// launchers (C++ linkage, used by rebop_gpu.cpp)

constexpr int kLdsCheckpoints = 32;

int lds_checkpoints() { return kLdsCheckpoints; }

// width: bytes per count (2, 4 or 8).  EXPR: the variant that also
// interprets rate expressions (jobs with Rate::Expr reactions; the others
// keep a kernel without the interpreter's stack); it keeps the records in
// LDS.  grec: records read in place (GREC).
template <bool EXPR, bool GREC>
static const void *lds_kernel_ptr_w(int width) {
  return width == 8   ? (const void *)&ssa_lds_kernel<kLdsCheckpoints, 8, EXPR, GREC>
         : width == 4 ? (const void *)&ssa_lds_kernel<kLdsCheckpoints, 4, EXPR, GREC>
                      : (const void *)&ssa_lds_kernel<kLdsCheckpoints, 2, EXPR, GREC>;
}

static const void *lds_kernel_ptr(int width, bool expr, bool grec) {
  if (expr) return lds_kernel_ptr_w<true, false>(width);
  return grec ? lds_kernel_ptr_w<false, true>(width) : lds_kernel_ptr_w<false, false>(width);
}

hipError_t lds_occupancy(int threads, size_t lds, int width, bool expr, bool grec,
                         int *blocks_per_cu) {
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu,
                                                      lds_kernel_ptr(width, expr, grec), threads,
                                                      lds);
}

hipError_t launch_lds(const rebop_dev_job *jobs, const rebop_dev_item *items, uint32_t n_items,
                      unsigned long long *counters, unsigned long long *job_events,
                      unsigned long long *job_queue, int blocks, int threads, size_t lds,
                      uint32_t rec_word_off, int width, bool expr, bool grec, bool rerun,
                      hipStream_t stream) {
  if (width != 2 && width != 4 && width != 8) return hipErrorInvalidValue;
  if (expr && grec) return hipErrorInvalidValue; // the interpreting kernel keeps them in LDS
  if (lds > 64 * 1024) {
    const hipError_t e = hipFuncSetAttribute(lds_kernel_ptr(width, expr, grec),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
#define REBOP_LDS_LAUNCH(W, E, G)                                                         \
  hipLaunchKernelGGL((ssa_lds_kernel<kLdsCheckpoints, W, E, G>), dim3(blocks), dim3(threads), \
                     lds, stream, jobs, items, n_items, counters, job_events, job_queue,     \
                     rec_word_off, rerun ? 1u : 0u)
  if (expr) {
    if (width == 8) REBOP_LDS_LAUNCH(8, true, false);
    else if (width == 4) REBOP_LDS_LAUNCH(4, true, false);
    else REBOP_LDS_LAUNCH(2, true, false);
  } else if (grec) {
    if (width == 8) REBOP_LDS_LAUNCH(8, false, true);
    else if (width == 4) REBOP_LDS_LAUNCH(4, false, true);
    else REBOP_LDS_LAUNCH(2, false, true);
  } else {
    if (width == 8) REBOP_LDS_LAUNCH(8, false, false);
    else if (width == 4) REBOP_LDS_LAUNCH(4, false, false);
    else REBOP_LDS_LAUNCH(2, false, false);
  }
#undef REBOP_LDS_LAUNCH
  return hipGetLastError();
}

} // namespace rebop
