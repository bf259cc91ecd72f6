/*
 * oracle.h -- CPU restatement of rebop's run-time direct-method SSA.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (rebop_amd/, include/rebop_gpu.h) never links or calls it.
 *
 * It restates, function by function (reference paths relative to the rebop
 * source tree):
 *   Rate::rate            src/gillespie.rs:71-90    orc_rate()
 *   Rate::sparse / dense  src/gillespie.rs:37-69    orc_net_add_reaction() packing
 *   Jump::affect          src/gillespie.rs:142-152  orc_affect()
 *   Gillespie::add_reaction asserts  :225-244       orc_net_add_reaction()
 *   Rate::expr / Expr::eval src/gillespie.rs:33-35, src/expr.rs:24-38
 *                                                   orc_net_add_reaction_expr()
 *   make_cumrates         src/gillespie.rs:357-364  orc_make_cumrates()
 *   choose_cumrate_sum    src/gillespie.rs:402-407  orc_choose_cumrate_sum()
 *   advance_until         src/gillespie.rs:315-344  orc_advance_until()
 *   _advance_one_reaction src/gillespie.rs:275-297  orc_advance_one_reaction()
 *   Gillespie.run grid    src/pyo3_gillespie.rs:200-208  orc_run_grid()
 *   Gillespie.run events  src/pyo3_gillespie.rs:209-223  orc_run_all_events()
 *   define_system!        src/gillespie_macro.rs:49-170 orc_net_new_system(),
 *                         orc_net_add_reaction_system(), orc_choose_first_above()
 * with the random source swapped for the Philox stream of rebop_rng.h
 * (rng kind ORC_RNG_PHILOX) or for an independent xoshiro256++ + libm log
 * stream (ORC_RNG_XOSHIRO) used for two-sample statistical tests and as the
 * second CPU-baseline variant.
 */
#ifndef REBOP_ORACLE_H
#define REBOP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_OK = 0, ORC_ERR_SHAPE = 1, ORC_ERR_INDEX = 2, ORC_ERR_VALUE = 3, ORC_ERR_ALLOC = 4 };
enum { ORC_RNG_PHILOX = 0, ORC_RNG_XOSHIRO = 1 };

typedef struct orc_net orc_net;

/* Gillespie::new(species, sparse) without the species vector (x0 is given per run). */
orc_net *orc_net_new(uint32_t nb_species, int sparse);
/* A define_system! network (gillespie_macro.rs:49-125): integer mass-action
 * products and the first-match choice; takes orc_net_add_reaction_system only. */
orc_net *orc_net_new_system(uint32_t nb_species);
void orc_net_free(orc_net *net);
uint32_t orc_net_nb_species(const orc_net *net);
uint32_t orc_net_nb_reactions(const orc_net *net);

/*
 * add_reaction(Rate::lma(k, orders), diffs) when rate_sparse == 0
 *   (orders has n entries, must equal nb_species), or
 * add_reaction(Rate::lma_sparse(k, [(idx[i], orders[i])]), diffs) when rate_sparse == 1
 *   (n pairs in the given order; every idx must be < nb_species).
 * diffs has ndiffs entries, must equal nb_species.  If the net was created
 * sparse, dense rates and the jump are converted as Rate::sparse / Jump::sparse do.
 */
int orc_net_add_reaction(orc_net *net, int rate_sparse, double k, const uint32_t *idx,
                         const uint32_t *orders, uint32_t n, const int64_t *diffs,
                         uint32_t ndiffs);

/*
 * add_reaction(Rate::expr(e), diffs): e as postfix tokens (rebop_expr.h:
 * op | arg << 8) over n_consts constants.  Its value at the counts (as f64)
 * is the propensity.  ORC_ERR_VALUE for a malformed program, ORC_ERR_INDEX for
 * an out-of-range constant or species index.
 */
int orc_net_add_reaction_expr(orc_net *net, const uint32_t *code, uint32_t n_code,
                              const double *consts, uint32_t n_consts, const int64_t *diffs,
                              uint32_t ndiffs);

/*
 * One define_system! reaction: `rate * lma(m_1 r_1) * ...` over the reactant
 * terms (r_idx, r_mult) in written order (repeats allowed), and the jump
 * r_i -= m_i, p_j += m_j.  ORC_ERR_VALUE on a net not made by
 * orc_net_new_system (whose other add_reaction calls return ORC_ERR_VALUE too).
 */
int orc_net_add_reaction_system(orc_net *net, double k, const uint32_t *r_idx,
                                const uint32_t *r_mult, uint32_t n_r, const uint32_t *p_idx,
                                const uint32_t *p_mult, uint32_t n_p);

/* Rate::rate of reaction r at counts x (for the known-answer tables). */
double orc_rate(const orc_net *net, uint32_t r, const int64_t *x);
/* make_cumrates: fills cum[R], returns the total. */
double orc_make_cumrates(const orc_net *net, const int64_t *x, double *cum);
/* choose_cumrate_sum over cum[0..n). */
uint64_t orc_choose_cumrate_sum(double chosen, const double *cum, uint64_t n);
/* _choice!: the first r < n with chosen < cum[r], else n. */
uint64_t orc_choose_first_above(double chosen, const double *cum, uint64_t n);
/* Jump::affect of reaction r on x (r >= R: no change). */
void orc_affect(const orc_net *net, uint32_t r, int64_t *x);

/* One trajectory's resumable state (Gillespie's species, t and rng). */
typedef struct {
  int64_t *x;      /* [nb_species], caller-owned */
  double t;
  uint64_t seed;
  uint64_t traj;
  uint64_t draw;   /* Philox: number of blocks drawn so far */
  uint64_t xs[4];  /* xoshiro256++ state */
  int rng_kind;
  uint64_t events; /* applied reactions */
  /* The device's watchdog, which the reference does not have: once an event
   * leaves draw >= draw_lim the state is treated as absorbing (capped = 1). */
  uint64_t draw_lim;
  int capped;
} orc_state;

void orc_state_init(orc_state *st, int64_t *x, uint64_t seed, uint64_t traj, int rng_kind);
/* Caps the draws from here on at `cap` (0: no cap). */
void orc_state_set_cap(orc_state *st, uint64_t cap);
void orc_advance_until(const orc_net *net, orc_state *st, double tmax, double *scratch_cum);
void orc_advance_one_reaction(const orc_net *net, orc_state *st, double *scratch_cum);

/*
 * The pyo3 grid loop for trajectories traj_begin .. traj_begin+n_traj-1:
 *   for i in 0..=nb_steps: t_i = tmax*i/nb_steps; advance_until(t_i); record.
 * x0: [nb_species] shared (x0_per_traj == 0) or [n_traj][nb_species].
 * out: [nb_steps+1][nb_species][n_traj] (trajectory fastest); may be NULL.
 * final_x (nullable): [n_traj][nb_species] counts after the last grid point.
 * times: [nb_steps+1] (nullable).  events (nullable): total applied reactions.
 * nthreads <= 1 runs on the calling thread; otherwise a pool hands out
 * chunks of trajectories from an atomic counter (64, or fewer for a small
 * call so that every thread gets at least four claims).
 */
int orc_run_grid(const orc_net *net, uint64_t seed, uint64_t traj_begin, uint64_t n_traj,
                 const int64_t *x0, int x0_per_traj, double tmax, uint64_t nb_steps,
                 int64_t *out, int64_t *final_x, double *times, uint64_t *events,
                 uint64_t *draws, int rng_kind, int nthreads, uint64_t draw_cap);

/* One network's grid run inside orc_run_grid_jobs (fields as orc_run_grid). */
typedef struct {
  const orc_net *net;
  uint64_t seed, traj_begin, n_traj;
  const int64_t *x0;
  int x0_per_traj;
  double tmax;
  uint64_t nb_steps;
  int64_t *out;     /* nullable */
  int64_t *final_x; /* nullable */
  uint64_t draw_cap; /* the device's watchdog (0: none), see orc_state */
} orc_grid_job;

/* Many grid runs (one per network) in one thread pool: the threads claim
 * chunks of trajectories of any job from one atomic counter, so a list of
 * small jobs keeps every thread busy.  *threads_busy (nullable) receives the
 * number of threads that ran at least one trajectory. */
int orc_run_grid_jobs(const orc_grid_job *jobs, uint32_t n_jobs, int rng_kind, int nthreads,
                      uint64_t *events, uint64_t *draws, int *threads_busy);

/*
 * The pyo3 nb_steps == 0 loop for one trajectory: records (t, x) before the
 * first event and after each event while t < tmax (an absorbing state sets
 * t = +inf).  Writes at most cap records; *n_out receives the number of
 * records the run produced (which may exceed cap: then the output is cut).
 * draw_cap (0: none) stops the loop after the event that reaches it, as the
 * device's watchdog does.
 */
int orc_run_all_events(const orc_net *net, uint64_t seed, uint64_t traj, const int64_t *x0,
                       double tmax, uint64_t cap, double *times, int64_t *xs,
                       uint64_t *n_out, int rng_kind, uint64_t draw_cap);

/* orc_run_all_events for trajectories traj_begin .. traj_begin + n_traj - 1
 * over nthreads threads (a timing harness: each thread writes the records
 * into a buffer of its own and drops them); *events and *records receive the
 * totals (events exclude the initial record and an absorbing state's
 * t = +inf record); *threads_busy (nullable): threads that ran a trajectory. */
int orc_run_all_events_many(const orc_net *net, uint64_t seed, uint64_t traj_begin, uint64_t n_traj,
                            const int64_t *x0, double tmax, uint64_t *events, uint64_t *records,
                            int rng_kind, int nthreads, int *threads_busy);

/* RNG primitives, exported so tests can pin them. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void orc_stream_block(uint64_t seed, uint64_t traj, uint64_t draw, uint32_t out[4]);
double orc_ln64(double x);
/* the rate-expression math of rebop_expr.h */
double orc_exp64(double x);
double orc_pow64(double x, double y);
double orc_fmax64(double a, double b);
double orc_fmin64(double a, double b);
double orc_u01_open_closed(uint32_t lo, uint32_t hi);
double orc_u01_closed_open(uint32_t lo, uint32_t hi);
double orc_grid_time(double tmax, uint64_t i, uint64_t nb_steps);

/* Number of hardware threads seen by the process. */
int orc_hw_threads(void);

#ifdef __cplusplus
}
#endif

#endif
