/*
 * ssa_layout.h -- device-side layout of a packed reaction network and of one
 * SSA launch, shared by the host library (rebop_gpu.cpp) and the kernels
 * (ssa_kernels.hip).
 *
 * Packed network (one contiguous device buffer, read by every lane at
 * wave-uniform addresses, so loads go through the scalar cache into SGPRs):
 *   k[R]            LMA rate constant of reaction r
 *   fac_begin[R+1]  factor list of reaction r is fac_begin[r] .. fac_begin[r+1]
 *   fac_s[F]        species index of each factor
 *   fac_c[F]        factor value is x[fac_s] - fac_c, as a double
 *   diff[R][S]      net stoichiometry, as doubles (staged into LDS per workgroup)
 * The factor list is Rate::rate's multiplication sequence (src/gillespie.rs:71-90):
 * for each (species s, order e) in evaluation order, the factors
 * (n+1-e), ..., n, i.e. c = e-1, ..., 0.  Evaluation order is ascending
 * species for a dense Rate::LMA and the stored pair order for Rate::LMASparse.
 *
 * The register and specialised kernels carry counts in FP64 registers.
 * Integers up to 2^53 are exact in binary64, so x - c, x + diff and the
 * products see exactly the values the reference's int64 counts produce (its
 * `x as f64` conversions are exact in that range).  The host rejects
 * |diff| > 2^20; these kernels check |x| <= 2^52 at every sample and that
 * fewer than 2^31 draws happened since the previous check (so no count can
 * have passed 2^53 in between), and raise REBOP_FLAG_EXACT otherwise: the
 * host then reruns the run on the LDS kernel's int64 counts, which check
 * every addition for int64 overflow (REBOP_FLAG_RANGE).  An infinite total
 * propensity (an overflowing product or rate expression) raises
 * REBOP_FLAG_RANGE too: time would not advance, so that trajectory is run
 * out as an absorbing state instead.
 */
#ifndef REBOP_SSA_LAYOUT_H
#define REBOP_SSA_LAYOUT_H

#include <stdint.h>

#define REBOP_EXACT_LIMIT 4503599627370496.0 /* 2^52 */
#define REBOP_DIFF_LIMIT (1ll << 20)
#define REBOP_DRAWS_PER_CHECK (1ull << 31)
/* the specialised kernels copy the sample times into LDS when nb_steps is
 * below this (at most 16 KB per workgroup); above it they read the table in
 * global memory (grouped kernels: compute them) */
#define REBOP_GRID_LDS 2048u

/* error flags raised by kernels */
#define REBOP_FLAG_RANGE 1u
#define REBOP_FLAG_INTERNAL 2u /* a launch invariant broke (reported as a device error) */
#define REBOP_FLAG_CAP 4u      /* a trajectory reached the draw cap (watchdog) */
#define REBOP_FLAG_NARROW 8u   /* a count left the int32 range of the narrow LDS kernel */
#define REBOP_FLAG_EXACT 16u   /* a count left the f64-exact range (|x| <= 2^52) of the
                                  register or specialised kernels: the host reruns such a
                                  run on the LDS kernel's int64 counts */

/* counters[] slots */
#define REBOP_CTR_EVENTS 0
#define REBOP_CTR_DRAWS 1
#define REBOP_CTR_TRAJ 2
#define REBOP_CTR_FLAGS 3
#define REBOP_CTR_QUEUE 4
#define REBOP_CTR_SPILL 5 /* trajectories handed from a persistent launch to its tail */
#define REBOP_CTR_CAPPED 6 /* trajectories stopped by the draw cap */
#define REBOP_CTR_JOBQ 7   /* trajectory queue of a single-job LDS launch */
#define REBOP_CTR_COUNT 8

typedef struct {
  uint32_t S, R, F, pad;
  const double *k;
  const uint32_t *fac_begin;
  const uint32_t *fac_s;
  const double *fac_c;
  const double *diff;
} rebop_dev_net;

/* modes of one launch */
#define REBOP_MODE_GRID 0   /* sample at tmax*i/nb_steps, i = 0..nb_steps */
#define REBOP_MODE_TARGET 1 /* one advance_until(tmax) from the stored state */
#define REBOP_MODE_EVENTS 2 /* every event while t < tmax (nb_steps == 0) */

typedef struct {
  uint64_t seed;
  uint64_t traj_begin;
  uint64_t n_traj;
  double tmax;
  uint64_t nb_steps;
  int32_t mode;
  int32_t persistent;
  /* Watchdog: a trajectory that has drawn draw_cap times in this launch (0 =
   * no cap) stops there as if its state were absorbing (no further draws;
   * remaining samples repeat its counts), raises REBOP_FLAG_CAP and is
   * counted in counters[REBOP_CTR_CAPPED]. */
  uint64_t draw_cap;
  const int64_t *x0;   /* [S] shared (x0_div == 0) or row j / x0_div of [*][S] */
  uint64_t x0_div;
  const double *rates; /* NULL, or row j / rates_div of [*][R]: per-trajectory rate constants */
  uint64_t rates_div;
  int64_t *out;        /* [nb+1][S][n] or NULL */
  int64_t *final_x;    /* [S][n] or NULL */
  double *st_x;        /* resumable state [S][n], or NULL */
  double *st_t;        /* [n] */
  uint64_t *st_draw;   /* [n] */
  unsigned long long *counters; /* [REBOP_CTR_COUNT] */
  unsigned long long *queue;    /* persistent launches: the next trajectory index to hand out
                                   (&counters[REBOP_CTR_QUEUE], or an LDS job's own queue) */
  const double *grid;  /* [nb_steps + 1] sample times tmax * i / nb_steps ([0] = tmax
                          in target mode), written before the launch; the LDS
                          kernel computes them instead */
  /* REBOP_MODE_EVENTS: records of trajectory j go to rows rec_offsets[j] ..
   * (row = time + counts); with rec_offsets == NULL only n_records[j] is set */
  uint64_t *n_records;         /* [n] */
  const uint64_t *rec_offsets; /* [n] or NULL */
  double *rec_times;           /* [total] */
  int64_t *rec_states;         /* [total][S] */
  /* Tail hand-off of a persistent launch.  Once the trajectory queue is
   * empty, a wave with at most spill_below lanes still running writes their
   * trajectories' state (index, grid position, time, stream position, range
   * mark, counts) to entry counters[REBOP_CTR_SPILL]++ of the spill arrays and
   * stops; a grouped launch with resume = 1 then continues entry e on the
   * lanes of group e, so the longest trajectories finish on several lanes
   * each instead of holding a mostly idle wave. */
  uint32_t spill_below; /* 0: no hand-off */
  uint32_t resume;      /* trajectories are spill entries */
  uint64_t spill_cap;   /* entries (and the row stride of spill_x) */
  uint64_t *spill_j;    /* [cap] local trajectory index */
  uint64_t *spill_draw; /* [cap] draws made */
  uint64_t *spill_mark; /* [cap] draw count at the last sample (range check) */
  double *spill_t;      /* [cap] */
  uint32_t *spill_i;    /* [cap] next grid point; bit 31: stopped by the draw cap */
  double *spill_x;      /* [S][cap] counts */
} rebop_dev_run;

/*
 * One network of an LDS-kernel launch (a "job"), read at wave-uniform
 * addresses.  Counts are int16, int32 or int64 in LDS (the launch's width);
 * reactions are grouped in blocks of `block` for the checkpointed cumulative
 * sum.  The jump is sparse.
 */
#define REBOP_JOB_LINEAR 1u /* cumulative sum may be non-monotone: full linear search */
#define REBOP_JOB_SLOW 2u   /* some reaction's factors are not all in its record */
#define REBOP_JOB_SYSTEM 4u /* define_system semantics: term products, first-match choice */
#define REBOP_LDS_LIMIT 1073741824 /* |count| <= 2^30 in the narrow (int32) LDS kernel */
#define REBOP_LDS16_LIMIT 32767    /* |count| <= 2^15 - 1 in the compact (int16) LDS kernel */

/*
 * Reaction records of the LDS kernel, 4 words per reaction, copied into LDS
 * when a wave starts an item of the job (read R times per event):
 *   w0, w1  rate constant k (f64 bits, low word first)
 *   w2, w3  the first two propensity factors (x[s] - c) in Rate::rate order:
 *           c in bits 0-11, bit 14 = present, the count row s + 1 in bits
 *           16-31 (one shift gives the row); an absent factor is 0 (row 0 of
 *           the counts holds ones, c = 0: it multiplies by exactly 1.0)
 *   w3      bit 12: the factors are not all here (more than two, or c >=
 *           4096): use fac_begin/fac_s/fac_c for this reaction
 *           bit 13: the jump is not all in the jump record (more than four
 *           entries): use jump_begin/jump_s/jump_d
 * and jump records, 4 words per reaction, read from global memory once per
 * event: up to four entries, species in bits 0-7, signed difference in bits
 * 8-31 (zero difference = unused).
 */
#define REBOP_REC_WORDS 4u
#define REBOP_REC_PRESENT 0x4000u
#define REBOP_JREC_WORDS 4u
#define REBOP_REC_SLOW_FACTORS 0x1000u
#define REBOP_REC_SLOW_JUMP 0x2000u

typedef struct {
  uint32_t S, R, block, nblocks;
  uint32_t flags, pad;
  const double *k;
  const uint32_t *fac_begin;
  const uint32_t *fac_s;
  const double *fac_c;
  const uint32_t *jump_begin; /* [R+1] */
  const uint32_t *jump_s;
  const int32_t *jump_d;
  const uint32_t *rec;  /* [R][REBOP_REC_WORDS] reaction records (copied to LDS) */
  const uint32_t *jrec; /* [R][REBOP_JREC_WORDS] jump records */
  /* the run of this job (fields as rebop_dev_run) */
  uint64_t seed;
  uint64_t traj_begin;
  uint64_t n_traj;
  double tmax;
  uint64_t nb_steps;
  int32_t mode;
  int32_t pad2;
  uint64_t draw_cap; /* as rebop_dev_run.draw_cap */
  uint64_t x0_div;
  const int64_t *x0;
  int64_t *out;
  int64_t *final_x;
  double *st_x;
  double *st_t;
  uint64_t *st_draw;
  uint64_t *n_records;
  const uint64_t *rec_offsets;
  double *rec_times;
  int64_t *rec_states;
  /* Rate::Expr reactions (NULL: the job has none): reaction r's postfix
   * program is expr_code[expr_begin[r] .. expr_begin[r + 1]) (rebop_expr.h;
   * empty for a mass-action reaction), its constant tokens index expr_const */
  const uint32_t *expr_begin; /* [R+1] */
  const uint32_t *expr_code;
  const double *expr_const;
  /* define_system reactions (NULL: not a define_system network): reaction
   * r's propensity is k[r] times the products lma(m, x[s]) over its reactant
   * terms sys_species / sys_mult [sys_begin[r] .. sys_begin[r + 1]) in
   * written order */
  const uint32_t *sys_begin; /* [R+1] */
  const uint32_t *sys_species;
  const uint32_t *sys_mult;
} rebop_dev_job;

/* a wave's unit of work: trajectories first .. first+count-1 of one job */
typedef struct {
  uint32_t job;
  uint32_t count;
  uint64_t first;
} rebop_dev_item;

#endif
