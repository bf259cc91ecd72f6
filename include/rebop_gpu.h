/*
 * rebop_gpu.h -- C ABI of the MI355X batched direct-method SSA.
 *
 * This is the drop-in boundary for rebop's run-time Gillespie path: the
 * entry points below are what a Rust FFI crate (or the Python extension)
 * binds in place of calling rebop's `gillespie::Gillespie` directly.  Each
 * declaration names the reference interface it replaces (paths relative to
 * the rebop source tree).  Plain C types only; every pointer argument is a
 * host pointer unless the function name ends in `_device`, in which case it
 * is a device pointer on the current HIP device.
 *
 * Semantics are those of rebop's `Gillespie::advance_until` driven by the
 * `Gillespie.run` grid loop, for N independent trajectories: trajectory j
 * (global id traj_begin + j) draws from the Philox4x32-10 stream keyed by
 * (seed, traj_begin + j); counts and grid samples are bit-identical to the
 * reference algorithm driven by that stream.
 *
 * Errors: every function returning int returns 0 (REBOP_OK) or one of the
 * REBOP_ERR_* codes; rebop_last_error() then describes the failure (the
 * message is thread-local and valid until the next failing call on that
 * thread).  Where the reference panics (`assert!`), the ABI returns an error.
 *
 * Streams: a handle keeps per-handle device state (launch counters, the
 * sample-time table, tail hand-off buffers), so the runs of one handle must
 * be issued on one stream at a time; distinct handles may run concurrently
 * on different streams.
 */
#ifndef REBOP_GPU_H
#define REBOP_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define REBOP_GPU_ABI_VERSION 2

enum rebop_status {
  REBOP_OK = 0,
  REBOP_ERR_SHAPE = 1,  /* length mismatch: assert_eq! at src/gillespie.rs:228, :237, :265 */
  REBOP_ERR_INDEX = 2,  /* species index out of range: assert! at src/gillespie.rs:229-233 */
  REBOP_ERR_VALUE = 3,  /* invalid argument value (e.g. nb_steps == 0, counts too large) */
  REBOP_ERR_DEVICE = 4, /* HIP runtime failure, or no usable gfx950 device */
  REBOP_ERR_ALLOC = 5,  /* host or device allocation failed */
  REBOP_ERR_RANGE = 6,  /* a count overflowed int64 (the reference's isize), or left
                           |x| <= 2^52 in a resumable batch or a run with rate tables,
                           or a total propensity was infinite */
  REBOP_ERR_UNSUPPORTED = 7, /* network shape beyond the compiled kernels */
  REBOP_ERR_CAP = 8          /* a trajectory reached the draw cap (rebop_gillespie_set_draw_cap) */
};

/* Counters of one run, summed over its trajectories. */
typedef struct rebop_stats {
  uint64_t events;       /* applied reactions (Jump::affect calls) */
  uint64_t draws;        /* Philox blocks drawn (events + discarded overshoot draws) */
  uint64_t trajectories; /* trajectories completed */
  uint64_t kernel;       /* which kernel ran: REBOP_KERNEL_* */
  double kernel_ms;      /* device time of the SSA kernel(s) (0 if not timed) */
  uint64_t tail;         /* trajectories finished by the grouped tail launch */
  uint64_t capped;       /* trajectories stopped by the draw cap (then REBOP_ERR_CAP) */
} rebop_stats;

enum rebop_kernel_kind {
  REBOP_KERNEL_AUTO = 0,
  REBOP_KERNEL_PLAIN = 1,      /* one trajectory per lane, lanes retire when done */
  REBOP_KERNEL_PERSISTENT = 2, /* lanes refill from an atomic trajectory queue */
  REBOP_KERNEL_LDS = 3,        /* counts in LDS: 256 bytes per species (int32; 512 with
                                  REBOP_KERNEL_WIDE, 128 with int16 counts) and 16 per
                                  reaction must fit one workgroup's LDS (160 KB) */
  /* flag: the network compiled into the kernel (run-time compilation,
   * cached); combine with PLAIN or PERSISTENT */
  REBOP_KERNEL_SPECIALIZED = 16,
  /* flags: 2 or 4 lanes per trajectory, sharing the Philox blocks and ln of
   * the next draws (PLAIN | SPECIALIZED grid runs; AUTO picks them when the
   * run has too few trajectories for more than one wave per SIMD) */
  REBOP_KERNEL_GROUP2 = 32,
  REBOP_KERNEL_GROUP4 = 64,
  /* flag: a persistent specialised grid or target run finishes every
   * trajectory in its one launch.  Without it, once the trajectory queue is
   * empty, waves with at most 16 lanes still running hand those trajectories
   * (state written to device memory) to a second launch that runs them with
   * 4 lanes each, which shortens the end of the run; results are identical */
  REBOP_KERNEL_WHOLE = 128,
  /* flag for the LDS kernel: counts as int64 (the reference's isize range)
   * instead of int32 (|x| <= 2^30; half the LDS).  Without it an LDS run
   * starts narrow unless the handle's initial counts are beyond int32, and a
   * run whose counters are read (stats non-NULL, or events_per_job) that
   * meets a count beyond int32 is rerun wide; resumable batches always run
   * wide (their state arrays are f64: |x| <= 2^52).
   *
   * The register and specialised kernels hold counts as f64, exact for
   * |x| <= 2^52.  A run with shared initial counts beyond that goes to the
   * wide LDS kernel up front; a run whose counts leave it (per-trajectory
   * initial counts, or growth) is rerun there when its counters are read
   * (stats non-NULL), with every output rewritten; with stats NULL the
   * launch is asynchronous and such a run's outputs are not exact. */
  REBOP_KERNEL_WIDE = 256,
  /* reported in rebop_stats.kernel: the LDS kernel ran with int16 counts
   * (|x| <= 2^15 - 1; half the LDS of int32).  An LDS run whose counters are
   * read starts there when the handle's initial counts fit it and the
   * smaller LDS lets more waves share a CU (large networks: otherwise the
   * registers bound the waves, and int32 runs as fast without the risk of a
   * rerun); a launch that meets a larger count stops early and is rerun
   * with int32 counts (then int64).  REBOP_LDS_COMPACT=0 in the environment
   * never starts at int16, =1 does whenever the initial counts fit. */
  REBOP_KERNEL_COMPACT = 512
};

/* Opaque network + initial state: rebop's `gillespie::Gillespie` (src/gillespie.rs:157-163). */
typedef struct rebop_gillespie rebop_gillespie;

/* Gillespie::new(species, sparse) -- src/gillespie.rs:168-176.  The seed is
 * drawn from the OS entropy source, as rand::make_rng() does. */
int rebop_gillespie_new(const int64_t *species, uint32_t nb_species, int sparse,
                        rebop_gillespie **out);

/* Gillespie::new_with_seed(species, sparse, seed) -- src/gillespie.rs:179-187. */
int rebop_gillespie_new_with_seed(const int64_t *species, uint32_t nb_species, int sparse,
                                  uint64_t seed, rebop_gillespie **out);

/* Drop for Gillespie.  Accepts NULL. */
void rebop_gillespie_free(rebop_gillespie *g);

/* Gillespie::seed(seed) -- src/gillespie.rs:189-191. */
int rebop_gillespie_seed(rebop_gillespie *g, uint64_t seed);

/* Watchdog, no counterpart in the reference (whose advance_until loop,
 * src/gillespie.rs:317-343, runs until tmax however many events that takes):
 * every later run of this handle stops a trajectory once it has made
 * max_draws random draws (events plus discarded overshoot draws) in that
 * call, treats its state as absorbing from there (no more draws; later
 * samples repeat its counts), and returns REBOP_ERR_CAP, with
 * rebop_stats.capped the number of such trajectories.  0 (the default)
 * means no cap, except for horizons that are not finite (tmax = +inf or
 * NaN, where only an absorbing state ends the reference's loop): those runs
 * use a cap of 2^20 draws unless the handle sets one.  The cap is checked
 * after each event, so a trajectory whose propensity keeps growing cannot
 * hold its wave, and the launch, indefinitely. */
int rebop_gillespie_set_draw_cap(rebop_gillespie *g, uint64_t max_draws);

/* Gillespie::nb_species / nb_reactions -- src/gillespie.rs:199-211. */
uint32_t rebop_gillespie_nb_species(const rebop_gillespie *g);
uint32_t rebop_gillespie_nb_reactions(const rebop_gillespie *g);

/* add_reaction(Rate::lma(rate, orders), differences) -- src/gillespie.rs:225-244
 * with Rate::lma at :24-26.  n_orders and n_differences must equal nb_species. */
int rebop_gillespie_add_reaction_lma(rebop_gillespie *g, double rate, const uint32_t *orders,
                                     uint32_t n_orders, const int64_t *differences,
                                     uint32_t n_differences);

/* add_reaction(Rate::lma_sparse(rate, [(species_idx[i], orders[i])]), differences)
 * -- src/gillespie.rs:225-244 with Rate::lma_sparse at :29-31.  The pairs are
 * evaluated in the given order (duplicates allowed), as Rate::rate does. */
int rebop_gillespie_add_reaction_lma_sparse(rebop_gillespie *g, double rate,
                                            const uint32_t *species_idx, const uint32_t *orders,
                                            uint32_t n_pairs, const int64_t *differences,
                                            uint32_t n_differences);

/* Rate expressions (Rate::Expr, src/expr.rs:9-38): the expression tree in
 * postfix order, one 32-bit token per node: op | (arg << 8), with
 * arg the constant index (CONST) or species index (SPECIES), 0 otherwise.
 * Evaluation follows Expr::eval (src/expr.rs:24-38): species counts as f64,
 * IEEE + - * / and negation, pow, max/min (a NaN operand yields the other
 * operand), exp; pow and exp are this library's own (host and device agree
 * bit for bit; the reference's libm may differ in the last ulp). */
#define REBOP_EXPR_CONST 1u
#define REBOP_EXPR_SPECIES 2u
#define REBOP_EXPR_NEG 3u
#define REBOP_EXPR_ADD 4u
#define REBOP_EXPR_SUB 5u
#define REBOP_EXPR_MUL 6u
#define REBOP_EXPR_DIV 7u
#define REBOP_EXPR_POW 8u
#define REBOP_EXPR_MAX 9u
#define REBOP_EXPR_MIN 10u
#define REBOP_EXPR_EXP 11u

/* add_reaction(Rate::expr(expr), differences) -- src/gillespie.rs:225-244 with
 * Rate::expr at :33-35.  The propensity is the expression's value (no
 * mass-action factor).  The program must be well formed (one value left, at
 * most 32 operands deep, at most 4096 tokens; REBOP_ERR_VALUE) and refer to
 * existing constants and species (REBOP_ERR_INDEX).  A network with rate
 * expressions runs on a network-specialised kernel, with the expressions
 * compiled in, when it fits one (S <= 32, R <= 64), and otherwise (or when
 * REBOP_KERNEL_LDS is asked for) on the LDS kernel, which interprets them;
 * fused job batches take such networks too.  The data-driven register
 * kernels never run them, and rate tables (rebop_run_sweep_device with
 * d_rates) return REBOP_ERR_UNSUPPORTED.  An expression that evaluates to
 * +inf makes the run return REBOP_ERR_RANGE (time would not advance). */
int rebop_gillespie_add_reaction_expr(rebop_gillespie *g, const uint32_t *code, uint32_t n_code,
                                      const double *consts, uint32_t n_consts,
                                      const int64_t *differences, uint32_t n_differences);

/*
 * define_system! networks -- src/gillespie_macro.rs:49-170.  The macro DSL's
 * arithmetic differs from Gillespie's: a reaction's propensity is
 * `rate * lma(m_1 r_1) * lma(m_2 r_2) * ...` over its reactant terms in written
 * order, where lma(m x) = x (x-1) ... (x-m+1) is an integer product (isize,
 * wrapping; m = 0 gives x, like the macro's loop) converted to f64 once
 * (_rate_lma!, :128-142), and the choice is the first reaction r with
 * `chosen < cum_r` (_choice!, :146-170; none: no reaction, time still
 * advances).  Draws and the time update are those of advance_until.
 *
 * rebop_system_new makes such a handle (rebop_gillespie_add_reaction_* then
 * return REBOP_ERR_UNSUPPORTED); every run entry point, batch and parameter
 * sweep accepts it.  It runs on a network-specialised kernel when the
 * network fits one (S <= 32, R <= 64), otherwise (or when REBOP_KERNEL_LDS
 * is asked for) on the LDS kernel, which also takes it in fused job batches;
 * never on the data-driven register kernels.
 */
#define REBOP_SYSTEM_MAX_COUNT 64u /* largest stoichiometric coefficient */
int rebop_system_new(const int64_t *species, uint32_t nb_species, uint64_t seed,
                     rebop_gillespie **out);
/* One reaction `m_1 r_1 + m_2 r_2 + ... => p_1 q_1 + ... @ rate`: reactants[i]
 * with coefficient reactant_counts[i] (repeats allowed: `A + A` is lma(1 A)^2,
 * not lma(2 A)), products likewise.  The rate is the value of the macro's
 * rate expression (finite, or NaN for a parameter never set: the system then
 * never reacts).  REBOP_ERR_INDEX for a species >= nb_species, REBOP_ERR_VALUE
 * for a coefficient above REBOP_SYSTEM_MAX_COUNT or a net change beyond 2^20. */
int rebop_system_add_reaction(rebop_gillespie *g, double rate, const uint32_t *reactants,
                              const uint32_t *reactant_counts, uint32_t n_reactants,
                              const uint32_t *products, const uint32_t *product_counts,
                              uint32_t n_products);

/* Uploads the network to the current device and compiles (cached) the
 * kernel `kernel` (REBOP_KERNEL_*; AUTO: the one an AUTO run would use) for
 * it, so that the first run does not pay for them.  Without a device it only
 * compiles, which checks the specialised kernel for gfx950.  No counterpart
 * in the reference (which has no device code to prepare). */
int rebop_gillespie_prepare(const rebop_gillespie *g, int kernel);

/* The autotuner's choice: which kernel and launch shape AUTO runs. */
typedef struct rebop_tune_result {
  int32_t kernel;   /* REBOP_KERNEL_* */
  int32_t threads;  /* workgroup size: 64, 128 or 256 */
  double occupancy; /* fraction of the resident workgroups a persistent launch uses */
  double ms;        /* pilot kernel time of that choice */
  uint32_t n_tried; /* variants timed */
} rebop_tune_result;

/* Times every applicable kernel variant (specialised / data-driven register,
 * persistent / plain, workgroup size 64/128/256, full or half resident grid,
 * LDS) on a pilot of n_pilot trajectories of this network (no output), and
 * makes REBOP_KERNEL_AUTO runs of this handle use the fastest.  All variants
 * give identical results, so the choice only affects speed.  The choice is
 * kept in an on-disk cache keyed by the network's shape (species, reactions,
 * total reactant order), nb_steps, the pilot size (power of two) and the
 * device; a later call for the same key reuses it without timing (n_tried =
 * 0).  REBOP_TUNE_CACHE names the file (empty: no cache; default under
 * $XDG_CACHE_HOME or ~/.cache), REBOP_TUNE_FORCE=1 re-times.  No counterpart
 * in the reference. */
int rebop_autotune(rebop_gillespie *g, double tmax, uint64_t nb_steps, uint64_t n_pilot,
                   rebop_tune_result *out);

/* Gillespie::set_species / get_species on the initial state -- src/gillespie.rs:260-267. */
int rebop_gillespie_set_species(rebop_gillespie *g, const int64_t *species, uint32_t n);
int rebop_gillespie_get_species(const rebop_gillespie *g, int64_t *species, uint32_t n);

/*
 * The grid loop of rebop's Python `Gillespie.run(tmax, nb_steps, seed=...)`
 * (src/pyo3_gillespie.rs:197-208: for i in 0..=nb_steps,
 * advance_until(tmax * i / nb_steps), record the counts), run for n_traj
 * independent trajectories on the GPU.
 *   x0      [n_traj][nb_species] per-trajectory initial counts, or NULL to
 *           start every trajectory from the handle's species vector.
 *   out     [nb_steps+1][nb_species][n_traj] int64, trajectory fastest (may be NULL).
 *   times   [nb_steps+1] sample times tmax*i/nb_steps (may be NULL).
 *   stats   counters of the run (may be NULL).
 * nb_steps must be > 0 (nb_steps == 0, every event, is rebop_run_all_events).
 * Host buffers; blocks until the results are copied.
 */
int rebop_run(const rebop_gillespie *g, double tmax, uint64_t nb_steps, uint64_t traj_begin,
              uint64_t n_traj, const int64_t *x0, int64_t *out, double *times,
              rebop_stats *stats);

/*
 * rebop_run with device-resident buffers, enqueued on `stream` (a
 * hipStream_t, NULL for the default stream).  d_x0 may be NULL; d_out may be
 * NULL (then only d_final and the counters are produced).  d_final
 * (nullable): [nb_species][n_traj] int64 counts after the last grid point.
 * kernel: REBOP_KERNEL_* (AUTO lets the tuner choose).  When stats is
 * non-NULL the call synchronises the stream to read the counters back and
 * to time the kernel; with stats == NULL it returns as soon as the work is
 * enqueued.
 */
int rebop_run_device(const rebop_gillespie *g, double tmax, uint64_t nb_steps,
                     uint64_t traj_begin, uint64_t n_traj, const int64_t *d_x0, int64_t *d_out,
                     int64_t *d_final, int kernel, void *stream, rebop_stats *stats);

/*
 * rebop_run_device without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory -- code
 * being captured into a HIP graph, tight serving loops, harnesses that forbid
 * hipMalloc.  Same arguments, results and errors as rebop_run_device, except:
 *   - the handle must be prepared on the current device first
 *     (rebop_gillespie_prepare with the kernel kind to be run; AUTO prepares
 *     what an AUTO run uses), else REBOP_ERR_VALUE;
 *   - the launch counters, sample-time table and work tables live in one
 *     fixed ~0.8 MB device region per GPU, which limits nb_steps to 65536
 *     and, for networks on the LDS kernel, n_traj to 2^20 (REBOP_ERR_VALUE
 *     beyond);
 *   - no tail hand-off (identical results) and stats->kernel_ms = 0.
 * All calls on one GPU share that region and it keeps the last call's
 * contents: calls that overlap on one GPU race, so issue them on one stream.
 */
int rebop_run_device_noalloc(const rebop_gillespie *g, double tmax, uint64_t nb_steps,
                             uint64_t traj_begin, uint64_t n_traj, const int64_t *d_x0,
                             int64_t *d_out, int64_t *d_final, int kernel, void *stream,
                             rebop_stats *stats);

/*
 * The nb_steps == 0 mode of rebop's Python `Gillespie.run`
 * (src/pyo3_gillespie.rs:209-223 over Gillespie::_advance_one_reaction,
 * src/gillespie.rs:275-297) for n_traj trajectories: the state before the
 * first event and after every event while t < tmax (the last event lies past
 * tmax; an absorbing state records t = +inf once).  Two calls: with times and
 * states NULL it fills n_records[j], trajectory j's record count (initial
 * state included); then, given those counts, it writes trajectory j's records
 * at row offset sum_{i<j} n_records[i] of times[] and states[][nb_species].
 * tmax may be +inf (every event until an absorbing state, as the reference's
 * loop does; a trajectory that does not reach one within the draw cap makes
 * the call return REBOP_ERR_CAP).  Host buffers.
 */
int rebop_run_all_events(const rebop_gillespie *g, double tmax, uint64_t traj_begin,
                         uint64_t n_traj, uint64_t *n_records, double *times, int64_t *states,
                         rebop_stats *stats);

/* rebop_run_all_events (src/pyo3_gillespie.rs:209-223 over
 * src/gillespie.rs:275-297) on device buffers, stream-ordered on `stream`
 * (hipStream_t, NULL = default), for callers whose records stay on the GPU.
 * Pass 1 (d_offsets, d_times, d_states NULL): d_n_records[j] = trajectory
 * j's record count.  Pass 2: d_offsets[j] = sum_{i<j} d_n_records[i] (the
 * caller's exclusive scan of pass 1's counts, which d_n_records must still
 * hold) and the records go to d_times[offset + r] and
 * d_states[offset + r][nb_species]; no trajectory writes more than
 * d_n_records[j] records, and a count that does not match its trajectory is
 * refused (REBOP_ERR_VALUE).  Same records, bit for bit, as the host
 * entry point.  Returns once the pass has finished. */
int rebop_run_all_events_device(const rebop_gillespie *g, double tmax, uint64_t traj_begin,
                                uint64_t n_traj, uint64_t *d_n_records, const uint64_t *d_offsets,
                                double *d_times, int64_t *d_states, void *stream,
                                rebop_stats *stats);

/* One network's run inside rebop_run_jobs_device. */
typedef struct rebop_job {
  const rebop_gillespie *net; /* network, initial species and seed */
  double tmax;
  uint64_t nb_steps;          /* > 0 */
  uint64_t traj_begin;        /* trajectory j uses the stream (seed, traj_begin + j) */
  uint64_t n_traj;
  const int64_t *d_x0;        /* device [n_traj][S], or NULL for the handle's species */
  int64_t *d_out;             /* device [nb_steps+1][S][n_traj], or NULL */
  int64_t *d_final;           /* device [S][n_traj], or NULL */
  double cost_hint;           /* expected relative work per trajectory (orders the
                                 work queue largest-first; 0 if unknown) */
} rebop_job;

/*
 * Many independent grid runs, one per job -- each is rebop_run_device's
 * src/pyo3_gillespie.rs:197-208 loop for its own network -- fused into one
 * launch of the LDS kernel: waves take (job, 64 trajectories) items from one
 * queue.  Results are identical to running the jobs one by one.  Enqueued on
 * `stream`; the call waits for completion only when stats or events_per_job
 * (nullable host array [n_jobs]: each job's applied-reaction count) is
 * given.  Its device tables are reused by later calls on the same stream.
 * No counterpart in the reference beyond calling `Gillespie.run` once per
 * network.
 */
int rebop_run_jobs_device(const rebop_job *jobs, uint32_t n_jobs, void *stream,
                          rebop_stats *stats, uint64_t *events_per_job);

/*
 * A parameter sweep over one network shape: n_points x traj_per_point
 * trajectories in one launch.  Trajectory q = p * traj_per_point + j (global
 * id traj_begin + q, stream (seed, traj_begin + q)) runs rebop_run_device's
 * grid loop with rate constants d_rates[p][0..R) (NULL: the network's own)
 * and initial counts d_x0[p][0..S) (NULL: the handle's species); d_out and
 * d_final are laid out as for rebop_run_device with n_traj = n_points *
 * traj_per_point.  Device pointers.  After rebop_gillespie_prepare the call
 * makes no host-device copy and no allocation, so with stats == NULL it can
 * be captured into a hipGraph and replayed (e.g. after rewriting d_rates).
 * Register and specialised kernels only (S <= 32, R <= 64).  The reference's
 * counterpart is one `Gillespie.run` per parameter point.
 */
int rebop_run_sweep_device(const rebop_gillespie *g, uint64_t n_points, uint64_t traj_per_point,
                           const double *d_rates, const int64_t *d_x0, double tmax,
                           uint64_t nb_steps, uint64_t traj_begin, int64_t *d_out,
                           int64_t *d_final, int kernel, void *stream, rebop_stats *stats);

/* Batched resumable state: N trajectories of one network, each with its own
 * (species, t, rng) -- the fields of gillespie::Gillespie (src/gillespie.rs:157-163)
 * for every trajectory.  Device-resident. */
typedef struct rebop_batch rebop_batch;

/* N copies of Gillespie::new_with_seed: trajectory j starts from x0[j] (or
 * the handle's species if x0 is NULL), t = 0, stream (seed, traj_begin + j). */
int rebop_batch_new(const rebop_gillespie *g, uint64_t traj_begin, uint64_t n_traj,
                    const int64_t *x0, rebop_batch **out);
void rebop_batch_free(rebop_batch *b);

/* Gillespie::advance_until(tmax) on every trajectory -- src/gillespie.rs:315-344. */
int rebop_batch_advance_until(rebop_batch *b, double tmax, rebop_stats *stats);

/* Gillespie::get_species (src/gillespie.rs:260-262) for every trajectory:
 * species [n_traj][nb_species]. */
int rebop_batch_get_species(const rebop_batch *b, int64_t *species);
/* Gillespie::set_species (src/gillespie.rs:264-267) for every trajectory. */
int rebop_batch_set_species(rebop_batch *b, const int64_t *species);
/* Gillespie::get_time / set_time (src/gillespie.rs:246-252) for every trajectory. */
int rebop_batch_get_time(const rebop_batch *b, double *t);
int rebop_batch_set_time(rebop_batch *b, const double *t);
/* The position of every trajectory's random stream (draws made so far; the
 * counter of its Philox stream).  With set_time and set_species it moves a
 * trajectory to another batch unchanged, as cloning a Gillespie (rng
 * included) does in the reference; no separate counterpart there. */
int rebop_batch_get_draws(const rebop_batch *b, uint64_t *draws);
int rebop_batch_set_draws(rebop_batch *b, const uint64_t *draws);

/* Description of the last error on this thread ("" if none). */
const char *rebop_last_error(void);

/* Name of the device the library runs on, and its compute-unit count. */
int rebop_device_info(char *name, size_t name_len, int *cu_count);

#ifdef __cplusplus
}
#endif

#endif /* REBOP_GPU_H */
