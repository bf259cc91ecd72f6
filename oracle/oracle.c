/*
 * oracle.c -- CPU restatement of rebop's run-time direct-method SSA.
 *
 * TEST INFRASTRUCTURE ONLY (the checker; see oracle.h).  Counts are int64
 * (Rust isize) and every count enters floating point through the same
 * int64 -> f64 conversion the reference performs, so this file follows the
 * reference's arithmetic literally; only the random source differs (the
 * Philox stream of rebop_rng.h, or xoshiro256++ for independent streams).
 *
 * Build: oracle/Makefile (gcc -O3 -ffp-contract=off; contraction would
 * change the rounding of the propensity products and the cumulative sum).
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "../rebop_amd/csrc/rebop_expr.h"
#include "../rebop_amd/csrc/rebop_rng.h"

/* ------------------------------------------------------------------ */
/* network: a list of (Rate, Jump) as Gillespie.reactions (gillespie.rs:160) */

struct orc_net {
  uint32_t S, R, cap;
  int sparse;         /* Gillespie.sparse (gillespie.rs:162) */
  int system;         /* define_system! semantics (gillespie_macro.rs:49-125) */
  /* Rate: LMA (dense, rate_sparse=0) holds S orders at rate_off[r]..+S;
   *       LMASparse holds (idx, order) pairs rate_off[r]..rate_off[r+1];
   *       a define_system! reaction (rate_sparse=3) holds its reactant terms
   *       (species, multiplicity) in written order at the same place. */
  int *rate_sparse;
  double *k;
  uint32_t *rate_off; /* R+1 */
  uint32_t *rate_idx, *rate_ord;
  uint32_t rate_len, rate_cap;
  /* Rate::Expr (rate_sparse[r] == 2): postfix tokens expr_off[r]..expr_off[r+1]
   * of expr_code; their constant indices are rebased into expr_const. */
  uint32_t *expr_off; /* R+1 */
  uint32_t *expr_code;
  double *expr_const;
  uint32_t expr_len, expr_cap, const_len, const_cap;
  /* Jump: Flat holds S diffs at jump_off[r]; Sparse holds (idx, diff) pairs. */
  int *jump_sparse;
  uint32_t *jump_off; /* R+1 */
  uint32_t *jump_idx;
  int64_t *jump_d;
  uint32_t jump_len, jump_cap;
};

orc_net *orc_net_new(uint32_t nb_species, int sparse) {
  orc_net *n = (orc_net *)calloc(1, sizeof *n);
  if (!n) return NULL;
  n->S = nb_species;
  n->sparse = sparse ? 1 : 0;
  n->rate_off = (uint32_t *)calloc(1, sizeof(uint32_t));
  n->jump_off = (uint32_t *)calloc(1, sizeof(uint32_t));
  n->expr_off = (uint32_t *)calloc(1, sizeof(uint32_t));
  if (!n->rate_off || !n->jump_off || !n->expr_off) {
    orc_net_free(n);
    return NULL;
  }
  return n;
}

orc_net *orc_net_new_system(uint32_t nb_species) {
  orc_net *n = orc_net_new(nb_species, 1);
  if (n) n->system = 1;
  return n;
}

void orc_net_free(orc_net *n) {
  if (!n) return;
  free(n->rate_sparse);
  free(n->k);
  free(n->rate_off);
  free(n->rate_idx);
  free(n->rate_ord);
  free(n->jump_sparse);
  free(n->jump_off);
  free(n->jump_idx);
  free(n->jump_d);
  free(n->expr_off);
  free(n->expr_code);
  free(n->expr_const);
  free(n);
}

uint32_t orc_net_nb_species(const orc_net *n) { return n->S; }
uint32_t orc_net_nb_reactions(const orc_net *n) { return n->R; }

static int grow(void **p, size_t elem, uint32_t *cap, uint32_t need) {
  if (need <= *cap) return 0;
  uint32_t nc = *cap ? *cap : 16;
  while (nc < need) nc *= 2;
  void *q = realloc(*p, (size_t)nc * elem);
  if (!q) return -1;
  *p = q;
  *cap = nc;
  return 0;
}

static int push_rate_entry(orc_net *n, uint32_t idx, uint32_t ord) {
  uint32_t c1 = n->rate_cap, c2 = n->rate_cap;
  if (grow((void **)&n->rate_idx, sizeof(uint32_t), &c1, n->rate_len + 1)) return -1;
  if (grow((void **)&n->rate_ord, sizeof(uint32_t), &c2, n->rate_len + 1)) return -1;
  n->rate_cap = c1;
  n->rate_idx[n->rate_len] = idx;
  n->rate_ord[n->rate_len] = ord;
  n->rate_len++;
  return 0;
}

static int push_jump_entry(orc_net *n, uint32_t idx, int64_t d) {
  uint32_t c1 = n->jump_cap, c2 = n->jump_cap;
  if (grow((void **)&n->jump_idx, sizeof(uint32_t), &c1, n->jump_len + 1)) return -1;
  if (grow((void **)&n->jump_d, sizeof(int64_t), &c2, n->jump_len + 1)) return -1;
  n->jump_cap = c1;
  n->jump_idx[n->jump_len] = idx;
  n->jump_d[n->jump_len] = d;
  n->jump_len++;
  return 0;
}

/* room for reaction r in every per-reaction array (the offsets need r + 2) */
static int reserve_reaction(orc_net *n, uint32_t r) {
  uint32_t c1 = n->cap, c2 = n->cap, c3 = n->cap, c4 = n->cap, c5 = n->cap, c6 = n->cap;
  if (grow((void **)&n->rate_sparse, sizeof(int), &c1, r + 2) ||
      grow((void **)&n->k, sizeof(double), &c2, r + 2) ||
      grow((void **)&n->jump_sparse, sizeof(int), &c3, r + 2) ||
      grow((void **)&n->rate_off, sizeof(uint32_t), &c4, r + 2) ||
      grow((void **)&n->jump_off, sizeof(uint32_t), &c5, r + 2) ||
      grow((void **)&n->expr_off, sizeof(uint32_t), &c6, r + 2))
    return -1;
  n->cap = c1;
  return 0;
}

static int push_jump(orc_net *n, uint32_t r, const int64_t *diffs, uint32_t ndiffs) {
  if (!n->sparse) {
    n->jump_sparse[r] = 0;
    for (uint32_t s = 0; s < ndiffs; ++s)
      if (push_jump_entry(n, s, diffs[s])) return ORC_ERR_ALLOC;
  } else {
    /* Jump::sparse: (index, difference) for difference != 0 */
    n->jump_sparse[r] = 1;
    for (uint32_t s = 0; s < ndiffs; ++s)
      if (diffs[s] != 0 && push_jump_entry(n, s, diffs[s])) return ORC_ERR_ALLOC;
  }
  n->jump_off[r + 1] = n->jump_len;
  return ORC_OK;
}

/* Gillespie::add_reaction (gillespie.rs:225-244) with Rate::sparse (:55-69)
 * and Jump::sparse (:126-140) applied when the net is sparse. */
int orc_net_add_reaction(orc_net *n, int rate_sparse, double k, const uint32_t *idx,
                         const uint32_t *orders, uint32_t cnt, const int64_t *diffs,
                         uint32_t ndiffs) {
  if (n->system) return ORC_ERR_VALUE;
  /* gillespie.rs:227-235: the rate must operate on existing species */
  if (!rate_sparse) {
    if (cnt != n->S) return ORC_ERR_SHAPE;
  } else {
    for (uint32_t i = 0; i < cnt; ++i)
      if (idx[i] >= n->S) return ORC_ERR_INDEX;
  }
  /* gillespie.rs:237: the jump must have one entry per species */
  if (ndiffs != n->S) return ORC_ERR_SHAPE;

  /* every per-reaction array is sized for R + 2 entries (the offsets need R + 1) */
  uint32_t r = n->R;
  if (r + 2 > n->cap) {
    if (reserve_reaction(n, r)) return ORC_ERR_ALLOC;
  }

  n->k[r] = k;
  if (!rate_sparse && !n->sparse) {
    /* Rate::LMA kept dense: one order per species, zeros included */
    n->rate_sparse[r] = 0;
    for (uint32_t s = 0; s < cnt; ++s)
      if (push_rate_entry(n, s, orders[s])) return ORC_ERR_ALLOC;
  } else if (!rate_sparse) {
    /* Rate::sparse: (index, exponent) for exponent > 0, ascending index */
    n->rate_sparse[r] = 1;
    for (uint32_t s = 0; s < cnt; ++s)
      if (orders[s] > 0 && push_rate_entry(n, s, orders[s])) return ORC_ERR_ALLOC;
  } else {
    /* Rate::LMASparse as given: order and duplicates preserved */
    n->rate_sparse[r] = 1;
    for (uint32_t i = 0; i < cnt; ++i)
      if (push_rate_entry(n, idx[i], orders[i])) return ORC_ERR_ALLOC;
  }
  n->rate_off[r + 1] = n->rate_len;
  n->expr_off[r + 1] = n->expr_len;
  int rc = push_jump(n, r, diffs, ndiffs);
  if (rc) return rc;
  n->R = r + 1;
  return ORC_OK;
}

/* Gillespie::add_reaction(Rate::expr(e), diffs) (gillespie.rs:225-244, :33-35):
 * Rate::sparse/dense leave an expression as it is (:51, :67); only the jump
 * follows the net's sparse flag.  The postfix program is checked for stack
 * discipline and in-range constant/species indices. */
int orc_net_add_reaction_expr(orc_net *n, const uint32_t *code, uint32_t n_code,
                              const double *consts, uint32_t n_consts, const int64_t *diffs,
                              uint32_t ndiffs) {
  if (n->system) return ORC_ERR_VALUE;
  if (ndiffs != n->S) return ORC_ERR_SHAPE;
  if (n_code == 0 || n_code > REBOP_EXPR_MAX_CODE) return ORC_ERR_VALUE;
  uint32_t depth = 0;
  for (uint32_t i = 0; i < n_code; ++i) {
    const uint32_t op = REBOP_EXPR_OPCODE(code[i]), arg = REBOP_EXPR_ARG(code[i]);
    if (op == REBOP_OP_CONST) {
      if (arg >= n_consts) return ORC_ERR_INDEX;
      depth++;
    } else if (op == REBOP_OP_SPECIES) {
      if (arg >= n->S) return ORC_ERR_INDEX;
      depth++;
    } else if (op == REBOP_OP_NEG || op == REBOP_OP_EXP) {
      if (depth < 1) return ORC_ERR_VALUE;
    } else if (op >= REBOP_OP_ADD && op <= REBOP_OP_MIN) {
      if (depth < 2) return ORC_ERR_VALUE;
      depth--;
    } else {
      return ORC_ERR_VALUE;
    }
    if (depth > REBOP_EXPR_MAX_DEPTH) return ORC_ERR_VALUE;
  }
  if (depth != 1) return ORC_ERR_VALUE;
  const uint32_t r = n->R;
  if (r + 2 > n->cap && reserve_reaction(n, r)) return ORC_ERR_ALLOC;
  uint32_t cc = n->expr_cap, kc = n->const_cap;
  if (grow((void **)&n->expr_code, sizeof(uint32_t), &cc, n->expr_len + n_code) ||
      grow((void **)&n->expr_const, sizeof(double), &kc, n->const_len + n_consts + 1))
    return ORC_ERR_ALLOC;
  n->expr_cap = cc;
  n->const_cap = kc;
  for (uint32_t i = 0; i < n_code; ++i) {
    uint32_t tok = code[i];
    if (REBOP_EXPR_OPCODE(tok) == REBOP_OP_CONST)
      tok = REBOP_EXPR_TOKEN(REBOP_OP_CONST, REBOP_EXPR_ARG(tok) + n->const_len);
    n->expr_code[n->expr_len++] = tok;
  }
  for (uint32_t i = 0; i < n_consts; ++i) n->expr_const[n->const_len++] = consts[i];
  n->rate_sparse[r] = 2;
  n->k[r] = 0.0;
  n->rate_off[r + 1] = n->rate_len;
  n->expr_off[r + 1] = n->expr_len;
  int rc = push_jump(n, r, diffs, ndiffs);
  if (rc) return rc;
  n->R = r + 1;
  return ORC_OK;
}

/* One define_system! reaction (gillespie_macro.rs:55-58, 106-112, 146-170):
 * `rate * lma(n_1 r_1) * lma(n_2 r_2) ...` over the reactant terms in written
 * order, then `r_i -= n_i` for every reactant term and `p_j += m_j` for every
 * product term; the jump is stored as the net difference per species (the
 * reference's separate integer decrements and increments give the same
 * counts).  Multiplicity 0 is accepted as the macro accepts the literal 0. */
int orc_net_add_reaction_system(orc_net *n, double k, const uint32_t *r_idx,
                                const uint32_t *r_mult, uint32_t n_r, const uint32_t *p_idx,
                                const uint32_t *p_mult, uint32_t n_p) {
  if (!n->system) return ORC_ERR_VALUE;
  for (uint32_t i = 0; i < n_r; ++i)
    if (r_idx[i] >= n->S) return ORC_ERR_INDEX;
  for (uint32_t i = 0; i < n_p; ++i)
    if (p_idx[i] >= n->S) return ORC_ERR_INDEX;
  const uint32_t r = n->R;
  if (r + 2 > n->cap && reserve_reaction(n, r)) return ORC_ERR_ALLOC;
  int64_t *d = (int64_t *)calloc(n->S ? n->S : 1, sizeof(int64_t));
  if (!d) return ORC_ERR_ALLOC;
  for (uint32_t i = 0; i < n_r; ++i) d[r_idx[i]] -= (int64_t)r_mult[i];
  for (uint32_t i = 0; i < n_p; ++i) d[p_idx[i]] += (int64_t)p_mult[i];
  n->k[r] = k;
  n->rate_sparse[r] = 3;
  for (uint32_t i = 0; i < n_r; ++i)
    if (push_rate_entry(n, r_idx[i], r_mult[i])) {
      free(d);
      return ORC_ERR_ALLOC;
    }
  n->rate_off[r + 1] = n->rate_len;
  n->expr_off[r + 1] = n->expr_len;
  int rc = push_jump(n, r, d, n->S);
  free(d);
  if (rc) return rc;
  n->R = r + 1;
  return ORC_OK;
}

/* _rate_lma! (gillespie_macro.rs:128-142): x (x-1) ... (x-m+1) in isize --
 * the loop `for i in 1..m { rate *= x - i }` starts from x, so m = 0 and
 * m = 1 both give x -- with release-build wrapping on overflow, then one
 * conversion to f64. */
static double system_lma(int64_t x, uint32_t m) {
  uint64_t p = (uint64_t)x;
  for (uint32_t i = 1; i < m; ++i) p = p * (uint64_t)(x - (int64_t)i);
  return (double)(int64_t)p;
}

/* Expr::eval (expr.rs:24-38) over the postfix program: counts enter as f64. */
static double eval_expr(const orc_net *net, uint32_t r, const int64_t *x) {
  double st[REBOP_EXPR_MAX_DEPTH];
  uint32_t sp = 0;
  for (uint32_t i = net->expr_off[r]; i < net->expr_off[r + 1]; ++i) {
    const uint32_t tok = net->expr_code[i];
    const uint32_t op = REBOP_EXPR_OPCODE(tok), arg = REBOP_EXPR_ARG(tok);
    switch (op) {
    case REBOP_OP_CONST: st[sp++] = net->expr_const[arg]; break;
    case REBOP_OP_SPECIES: st[sp++] = (double)x[arg]; break;
    case REBOP_OP_NEG: st[sp - 1] = -st[sp - 1]; break;
    case REBOP_OP_EXP: st[sp - 1] = rebop_exp64(st[sp - 1]); break;
    default: {
      const double b = st[--sp], a = st[sp - 1];
      double v;
      switch (op) {
      case REBOP_OP_ADD: v = a + b; break;
      case REBOP_OP_SUB: v = a - b; break;
      case REBOP_OP_MUL: v = a * b; break;
      case REBOP_OP_DIV: v = a / b; break;
      case REBOP_OP_POW: v = rebop_pow64(a, b); break;
      case REBOP_OP_MAX: v = rebop_fmax64(a, b); break;
      default: v = rebop_fmin64(a, b); break;
      }
      st[sp - 1] = v;
    }
    }
  }
  return st[0];
}

/* ------------------------------------------------------------------ */
/* Rate::rate (gillespie.rs:71-90).  Dense: fold over species ascending,
 * each (n+1-e ..= n) multiplied in ascending order as f64.  Sparse: the same
 * over the stored pairs, in stored order.  Both are the same loop here since
 * the dense form stores every species in ascending order. */
double orc_rate(const orc_net *net, uint32_t r, const int64_t *x) {
  if (net->rate_sparse[r] == 2) return eval_expr(net, r, x);
  if (net->rate_sparse[r] == 3) {
    /* gillespie_macro.rs:108: `$rate * lma(..) * lma(..)`, left to right */
    double a = net->k[r];
    for (uint32_t i = net->rate_off[r]; i < net->rate_off[r + 1]; ++i)
      a = a * system_lma(x[net->rate_idx[i]], net->rate_ord[i]);
    return a;
  }
  double acc = net->k[r];
  for (uint32_t i = net->rate_off[r]; i < net->rate_off[r + 1]; ++i) {
    const int64_t n = x[net->rate_idx[i]];
    const int64_t e = (int64_t)net->rate_ord[i];
    for (int64_t f = n + 1 - e; f <= n; ++f) acc = acc * (double)f;
  }
  return acc;
}

/* Jump::affect (gillespie.rs:142-152); r == R (a define_system! choice that
 * matched no reaction) changes nothing */
void orc_affect(const orc_net *net, uint32_t r, int64_t *x) {
  if (r >= net->R) return;
  for (uint32_t i = net->jump_off[r]; i < net->jump_off[r + 1]; ++i)
    x[net->jump_idx[i]] += net->jump_d[i];
}

/* make_cumrates (gillespie.rs:357-364): cum[r] = total + rate_r; total = cum[r] */
double orc_make_cumrates(const orc_net *net, const int64_t *x, double *cum) {
  double total = 0.0;
  for (uint32_t r = 0; r < net->R; ++r) {
    cum[r] = total + orc_rate(net, r, x);
    total = cum[r];
  }
  return total;
}

/* choose_cumrate_sum (gillespie.rs:402-407): #{r : cum[r] < chosen} */
uint64_t orc_choose_cumrate_sum(double chosen, const double *cum, uint64_t n) {
  uint64_t k = 0;
  for (uint64_t r = 0; r < n; ++r) k += (cum[r] < chosen) ? 1u : 0u;
  return k;
}

/* _choice! (gillespie_macro.rs:146-170): the first r with chosen < carry +
 * rate_r, where carry is `0. + rate_0 + ... + rate_{r-1}` -- the same sum as
 * cum[r] -- and n (no reaction applied) when no r qualifies. */
uint64_t orc_choose_first_above(double chosen, const double *cum, uint64_t n) {
  for (uint64_t r = 0; r < n; ++r)
    if (chosen < cum[r]) return r;
  return n;
}

static inline uint64_t choose(const orc_net *net, double chosen, const double *cum) {
  return net->system ? orc_choose_first_above(chosen, cum, net->R)
                     : orc_choose_cumrate_sum(chosen, cum, net->R);
}

/* ------------------------------------------------------------------ */
/* random sources */

static inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static uint64_t splitmix64(uint64_t *s) {
  uint64_t z = (*s += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

static uint64_t xoshiro_next(uint64_t *s) {
  const uint64_t result = rotl64(s[0] + s[3], 23) + s[0];
  const uint64_t t = s[1] << 17;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = rotl64(s[3], 45);
  return result;
}

void orc_state_init(orc_state *st, int64_t *x, uint64_t seed, uint64_t traj, int rng_kind) {
  st->x = x;
  st->t = 0.0;
  st->seed = seed;
  st->traj = traj;
  st->draw = 0;
  st->rng_kind = rng_kind;
  st->events = 0;
  st->draw_lim = UINT64_MAX;
  st->capped = 0;
  uint64_t sm = seed ^ (traj * 0xd1342543de82ef95ull) ^ 0x5851f42d4c957f2dull;
  for (int i = 0; i < 4; ++i) st->xs[i] = splitmix64(&sm);
}

void orc_state_set_cap(orc_state *st, uint64_t cap) {
  st->draw_lim = (cap == 0 || st->draw + cap < st->draw) ? UINT64_MAX : st->draw + cap;
}

/* one draw: the exponential, and (for Philox) the uniform of the same block */
static inline double draw_exp1(orc_state *st, double *u_out) {
  if (st->rng_kind == ORC_RNG_PHILOX) {
    const rebop_u32x4 b = rebop_draw_block(st->seed, st->traj, st->draw);
    st->draw++;
    *u_out = rebop_unif_of(b);
    return rebop_exp1_of(b);
  }
  const uint64_t a = xoshiro_next(st->xs);
  const uint64_t c = xoshiro_next(st->xs);
  st->draw++;
  *u_out = (double)(c >> 11) * 0x1p-53;
  return -log((double)((a >> 11) + 1) * 0x1p-53);
}

/* Gillespie::advance_until (gillespie.rs:315-344); for a define_system! net
 * the macro's advance_until (gillespie_macro.rs:99-124), which differs only in
 * the propensity products (orc_rate) and the choice (choose) */
void orc_advance_until(const orc_net *net, orc_state *st, double tmax, double *cum) {
  for (;;) {
    const double total = orc_make_cumrates(net, st->x, cum);
    if (!(0.0 < total) || st->capped) { /* :323-326 absorbing: no draw */
      st->t = tmax;
      return;
    }
    double u;
    const double e = draw_exp1(st, &u);
    st->t = st->t + e / total; /* :327 */
    if (st->t > tmax) {        /* :328-331 the draw is consumed */
      st->t = tmax;
      return;
    }
    const double chosen = total * u; /* :332 */
    const uint64_t k = choose(net, chosen, cum);
    orc_affect(net, (uint32_t)k, st->x); /* :342 */
    st->events++;
    if (st->draw >= st->draw_lim) st->capped = 1; /* the device's watchdog */
  }
}

/* Gillespie::_advance_one_reaction (gillespie.rs:275-297) */
void orc_advance_one_reaction(const orc_net *net, orc_state *st, double *cum) {
  const double total = orc_make_cumrates(net, st->x, cum);
  if (!(0.0 < total)) {
    st->t = INFINITY;
    return;
  }
  double u;
  const double e = draw_exp1(st, &u);
  st->t = st->t + e / total;
  const double chosen = total * u;
  const uint64_t k = choose(net, chosen, cum);
  orc_affect(net, (uint32_t)k, st->x);
  st->events++;
  if (st->draw >= st->draw_lim) st->capped = 1; /* the device's watchdog */
}

/* pyo3_gillespie.rs:202 */
double orc_grid_time(double tmax, uint64_t i, uint64_t nb_steps) {
  return tmax * (double)i / (double)nb_steps;
}

/* ------------------------------------------------------------------ */
/* the grid driver, pooled over threads */

/* Trajectories a worker takes per atomic claim: at most 64, fewer when the
 * call is small, so that every thread gets work (four claims per thread at
 * least, down to one trajectory per claim). */
static uint64_t pool_chunk(uint64_t total, int nthreads) {
  const uint64_t per = nthreads > 1 ? (uint64_t)nthreads * 4u : 1u;
  uint64_t c = (total + per - 1) / per;
  if (c < 1) c = 1;
  if (c > 64) c = 64;
  return c;
}

typedef struct {
  const orc_grid_job *jobs;
  uint32_t n_jobs;
  const uint64_t *chunk_end; /* [n_jobs] prefix sums of each job's chunk count */
  uint64_t chunk;
  uint32_t max_s, max_r;
  int rng_kind;
  atomic_uint_fast64_t next;
  atomic_uint_fast64_t events;
  atomic_uint_fast64_t draws;
  atomic_int busy;
  atomic_int err;
} grid_pool;

/* One trajectory of a job: the pyo3 grid loop (pyo3_gillespie.rs:200-208). */
static void grid_one(const orc_grid_job *J, uint64_t j, int rng_kind, int64_t *x, double *cum,
                     uint64_t *ev, uint64_t *dr) {
  const uint32_t S = J->net->S;
  const int64_t *src = J->x0_per_traj ? J->x0 + j * S : J->x0;
  memcpy(x, src, sizeof(int64_t) * S);
  orc_state st;
  orc_state_init(&st, x, J->seed, J->traj_begin + j, rng_kind);
  orc_state_set_cap(&st, J->draw_cap);
  for (uint64_t i = 0; i <= J->nb_steps; ++i) {
    const double t = orc_grid_time(J->tmax, i, J->nb_steps);
    orc_advance_until(J->net, &st, t, cum);
    if (J->out)
      for (uint32_t s = 0; s < S; ++s) J->out[(i * S + s) * J->n_traj + j] = x[s];
  }
  if (J->final_x) memcpy(J->final_x + j * S, x, sizeof(int64_t) * S);
  *ev += st.events;
  *dr += st.draw;
}

static void *grid_worker(void *arg) {
  grid_pool *P = (grid_pool *)arg;
  int64_t *x = (int64_t *)malloc(sizeof(int64_t) * (P->max_s ? P->max_s : 1));
  double *cum = (double *)malloc(sizeof(double) * (P->max_r ? P->max_r : 1));
  if (!x || !cum) {
    atomic_store(&P->err, ORC_ERR_ALLOC);
    free(x);
    free(cum);
    return NULL;
  }
  const uint64_t n_chunks = P->n_jobs ? P->chunk_end[P->n_jobs - 1] : 0;
  uint64_t ev = 0, dr = 0;
  int ran = 0;
  uint32_t jb = 0; /* claims only grow: the job index only moves forward */
  for (;;) {
    const uint64_t c = atomic_fetch_add(&P->next, 1);
    if (c >= n_chunks) break;
    while (P->chunk_end[jb] <= c) jb++;
    const orc_grid_job *J = &P->jobs[jb];
    const uint64_t c0 = jb ? P->chunk_end[jb - 1] : 0;
    const uint64_t b = (c - c0) * P->chunk;
    const uint64_t e = b + P->chunk < J->n_traj ? b + P->chunk : J->n_traj;
    for (uint64_t j = b; j < e; ++j) grid_one(J, j, P->rng_kind, x, cum, &ev, &dr);
    ran = 1;
  }
  atomic_fetch_add(&P->events, ev);
  atomic_fetch_add(&P->draws, dr);
  if (ran) atomic_fetch_add(&P->busy, 1);
  free(x);
  free(cum);
  return NULL;
}

int orc_run_grid_jobs(const orc_grid_job *jobs, uint32_t n_jobs, int rng_kind, int nthreads,
                      uint64_t *events, uint64_t *draws, int *threads_busy) {
  grid_pool P;
  uint64_t total = 0;
  for (uint32_t q = 0; q < n_jobs; ++q) {
    if (jobs[q].nb_steps == 0) return ORC_ERR_VALUE;
    total += jobs[q].n_traj;
  }
  P.jobs = jobs;
  P.n_jobs = n_jobs;
  P.chunk = pool_chunk(total, nthreads);
  P.rng_kind = rng_kind;
  P.max_s = 0;
  P.max_r = 0;
  uint64_t *ends = (uint64_t *)malloc(sizeof(uint64_t) * (n_jobs ? n_jobs : 1));
  if (!ends) return ORC_ERR_ALLOC;
  uint64_t acc = 0;
  for (uint32_t q = 0; q < n_jobs; ++q) {
    acc += (jobs[q].n_traj + P.chunk - 1) / P.chunk;
    ends[q] = acc;
    if (jobs[q].net->S > P.max_s) P.max_s = jobs[q].net->S;
    if (jobs[q].net->R > P.max_r) P.max_r = jobs[q].net->R;
  }
  P.chunk_end = ends;
  atomic_init(&P.next, 0);
  atomic_init(&P.events, 0);
  atomic_init(&P.draws, 0);
  atomic_init(&P.busy, 0);
  atomic_init(&P.err, 0);
  if (nthreads <= 1) {
    grid_worker(&P);
  } else {
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
    if (!th) {
      free(ends);
      return ORC_ERR_ALLOC;
    }
    int started = 0;
    for (int i = 0; i < nthreads; ++i) {
      if (pthread_create(&th[i], NULL, grid_worker, &P) != 0) break;
      started++;
    }
    if (started == 0) grid_worker(&P);
    for (int i = 0; i < started; ++i) pthread_join(th[i], NULL);
    free(th);
  }
  free(ends);
  if (events) *events = atomic_load(&P.events);
  if (draws) *draws = atomic_load(&P.draws);
  if (threads_busy) *threads_busy = atomic_load(&P.busy);
  return atomic_load(&P.err);
}

int orc_run_grid(const orc_net *net, uint64_t seed, uint64_t traj_begin, uint64_t n_traj,
                 const int64_t *x0, int x0_per_traj, double tmax, uint64_t nb_steps,
                 int64_t *out, int64_t *final_x, double *times, uint64_t *events,
                 uint64_t *draws, int rng_kind, int nthreads, uint64_t draw_cap) {
  if (nb_steps == 0) return ORC_ERR_VALUE;
  if (times)
    for (uint64_t i = 0; i <= nb_steps; ++i) times[i] = orc_grid_time(tmax, i, nb_steps);
  orc_grid_job J;
  J.net = net;
  J.seed = seed;
  J.traj_begin = traj_begin;
  J.n_traj = n_traj;
  J.x0 = x0;
  J.x0_per_traj = x0_per_traj;
  J.tmax = tmax;
  J.nb_steps = nb_steps;
  J.out = out;
  J.final_x = final_x;
  J.draw_cap = draw_cap;
  return orc_run_grid_jobs(&J, 1, rng_kind, nthreads, events, draws, NULL);
}

/* pyo3_gillespie.rs:209-223 for one trajectory */
int orc_run_all_events(const orc_net *net, uint64_t seed, uint64_t traj, const int64_t *x0,
                       double tmax, uint64_t cap, double *times, int64_t *xs,
                       uint64_t *n_out, int rng_kind, uint64_t draw_cap) {
  const uint32_t S = net->S;
  int64_t *x = (int64_t *)malloc(sizeof(int64_t) * (S ? S : 1));
  double *cum = (double *)malloc(sizeof(double) * (net->R ? net->R : 1));
  if (!x || !cum) {
    free(x);
    free(cum);
    return ORC_ERR_ALLOC;
  }
  memcpy(x, x0, sizeof(int64_t) * S);
  orc_state st;
  orc_state_init(&st, x, seed, traj, rng_kind);
  orc_state_set_cap(&st, draw_cap);
  uint64_t n = 0;
#define REC()                                                        \
  do {                                                               \
    if (n < cap) {                                                   \
      times[n] = st.t;                                               \
      for (uint32_t s = 0; s < S; ++s) xs[n * S + s] = x[s];         \
    }                                                                \
    n++;                                                             \
  } while (0)
  REC();
  while (st.t < tmax && !st.capped) {
    orc_advance_one_reaction(net, &st, cum);
    REC();
  }
#undef REC
  *n_out = n;
  free(x);
  free(cum);
  return ORC_OK;
}

/* orc_run_all_events over n_traj trajectories, pooled over threads (blocks
 * of pool_chunk trajectories): each thread records its trajectories into a buffer of its
 * own (the records are written, then dropped: a timing harness); events and
 * records are summed. */
typedef struct {
  const orc_net *net;
  uint64_t seed, traj_begin, n_traj;
  const int64_t *x0;
  double tmax;
  int rng_kind;
  uint64_t chunk;
  atomic_uint_fast64_t next;
  atomic_uint_fast64_t events;
  atomic_uint_fast64_t records;
  atomic_int busy;
  atomic_int err;
} events_job;

static void *events_worker(void *arg) {
  events_job *J = (events_job *)arg;
  const uint32_t S = J->net->S;
  uint64_t cap = 4096;
  double *times = (double *)malloc(sizeof(double) * cap);
  int64_t *xs = (int64_t *)malloc(sizeof(int64_t) * cap * (S ? S : 1));
  uint64_t ev = 0, recs = 0;
  int ran = 0;
  if (!times || !xs) {
    atomic_store(&J->err, ORC_ERR_ALLOC);
    free(times);
    free(xs);
    return NULL;
  }
  for (;;) {
    const uint64_t b = atomic_fetch_add(&J->next, J->chunk);
    if (b >= J->n_traj) break;
    const uint64_t e = b + J->chunk < J->n_traj ? b + J->chunk : J->n_traj;
    ran = 1;
    for (uint64_t j = b; j < e; ++j) {
      uint64_t n = 0;
      int rc = orc_run_all_events(J->net, J->seed, J->traj_begin + j, J->x0, J->tmax, cap, times, xs,
                                  &n, J->rng_kind, 0);
      if (rc == ORC_OK && n > cap) { /* longer than the buffer: grow it, run again */
        free(times);
        free(xs);
        cap = n;
        times = (double *)malloc(sizeof(double) * cap);
        xs = (int64_t *)malloc(sizeof(int64_t) * cap * (S ? S : 1));
        if (!times || !xs) rc = ORC_ERR_ALLOC;
        else rc = orc_run_all_events(J->net, J->seed, J->traj_begin + j, J->x0, J->tmax, cap, times, xs,
                                     &n, J->rng_kind, 0);
      }
      if (rc != ORC_OK) {
        atomic_store(&J->err, rc);
        break;
      }
      /* records = the initial state, one per event, and one t = +inf record
       * for an absorbing state */
      ev += n - 1 - (isinf(times[n - 1]) ? 1 : 0);
      recs += n;
    }
  }
  atomic_fetch_add(&J->events, ev);
  atomic_fetch_add(&J->records, recs);
  if (ran) atomic_fetch_add(&J->busy, 1);
  free(times);
  free(xs);
  return NULL;
}

int orc_run_all_events_many(const orc_net *net, uint64_t seed, uint64_t traj_begin, uint64_t n_traj,
                            const int64_t *x0, double tmax, uint64_t *events, uint64_t *records,
                            int rng_kind, int nthreads, int *threads_busy) {
  events_job J;
  J.chunk = pool_chunk(n_traj, nthreads);
  atomic_init(&J.busy, 0);
  J.net = net;
  J.seed = seed;
  J.traj_begin = traj_begin;
  J.n_traj = n_traj;
  J.x0 = x0;
  J.tmax = tmax;
  J.rng_kind = rng_kind;
  atomic_init(&J.next, 0);
  atomic_init(&J.events, 0);
  atomic_init(&J.records, 0);
  atomic_init(&J.err, 0);
  if (nthreads <= 1) {
    events_worker(&J);
  } else {
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
    if (!th) return ORC_ERR_ALLOC;
    int started = 0;
    for (int i = 0; i < nthreads; ++i) {
      if (pthread_create(&th[i], NULL, events_worker, &J) != 0) break;
      started++;
    }
    if (started == 0) events_worker(&J);
    for (int i = 0; i < started; ++i) pthread_join(th[i], NULL);
    free(th);
  }
  if (events) *events = atomic_load(&J.events);
  if (records) *records = atomic_load(&J.records);
  if (threads_busy) *threads_busy = atomic_load(&J.busy);
  return atomic_load(&J.err);
}

/* ------------------------------------------------------------------ */

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  const rebop_u32x4 r = rebop_philox4x32_10(ctr[0], ctr[1], ctr[2], ctr[3], key[0], key[1]);
  for (int i = 0; i < 4; ++i) out[i] = r.w[i];
}

/* the per-trajectory form of the same block that the device kernels draw with */
void orc_stream_block(uint64_t seed, uint64_t traj, uint64_t draw, uint32_t out[4]) {
  const rebop_u32x4 r = rebop_stream_block(rebop_stream_of(seed, traj), draw);
  for (int i = 0; i < 4; ++i) out[i] = r.w[i];
}

double orc_ln64(double x) { return rebop_ln64(x); }
double orc_exp64(double x) { return rebop_exp64(x); }
double orc_pow64(double x, double y) { return rebop_pow64(x, y); }
double orc_fmax64(double a, double b) { return rebop_fmax64(a, b); }
double orc_fmin64(double a, double b) { return rebop_fmin64(a, b); }
double orc_u01_open_closed(uint32_t lo, uint32_t hi) { return rebop_u01_open_closed(lo, hi); }
double orc_u01_closed_open(uint32_t lo, uint32_t hi) { return rebop_u01_closed_open(lo, hi); }

int orc_hw_threads(void) {
  long n = sysconf(_SC_NPROCESSORS_ONLN);
  return n > 0 ? (int)n : 1;
}
