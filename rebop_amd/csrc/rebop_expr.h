/*
 * rebop_expr.h -- rate expressions (the reference's Rate::Expr) as postfix
 * bytecode, and the exp / pow / max / min they evaluate with.
 *
 * Compiled by gcc (host, C11: the oracle's interpreter) and by hipcc / hipRTC
 * (gfx950: the network-specialised kernels inline the same functions), so an
 * expression gives bit-identical propensities on both sides.
 *
 * Reference: src/expr.rs:9-38 (enum Expr and Expr::eval):
 *   Constant(c) -> c;  Concentration(i) -> species[i] as f64;  Neg, Add, Sub,
 *   Mul, Div -> IEEE ops;  Pow -> a.powf(b);  Max -> a.max(b);  Min -> a.min(b);
 *   Exp -> a.exp().
 * The tree is flattened to postfix: one 32-bit token per node, low 8 bits the
 * operation, high 24 bits its argument (constant index or species index).
 *
 * exp / pow are written here rather than taken from a math library so that host
 * and device agree bit for bit (the reference's powf / exp come from the Rust
 * toolchain's libm; their last-ulp behaviour is not reproduced):
 *  - rebop_exp64: Cody-Waite reduction x = k ln2 + r with ln 2 split as in
 *    rebop_rng.h, |r| <= ln2/2, exp(r) by its Taylor series to degree 13
 *    (truncation < 2^-57 relative), Horner with fma, then scaling by 2^k built
 *    from the exponent bits (two factors, so subnormal results round once).
 *  - rebop_pow64: the special cases of C99 Annex F.9.4.4 (which Rust's powf
 *    follows), x^n by binary powering for integral |n| <= 64 (exact whenever
 *    the result and every partial product are representable), and
 *    exp(y ln x) otherwise.
 *  - max / min: IEEE maxNum / minNum as in Rust's f64::max / f64::min (a NaN
 *    operand yields the other operand); equal operands return the second.
 */
#ifndef REBOP_EXPR_H
#define REBOP_EXPR_H

#include "rebop_rng.h"

/* bytecode operations (same values as REBOP_EXPR_* in the public header) */
#define REBOP_OP_CONST 1u
#define REBOP_OP_SPECIES 2u
#define REBOP_OP_NEG 3u
#define REBOP_OP_ADD 4u
#define REBOP_OP_SUB 5u
#define REBOP_OP_MUL 6u
#define REBOP_OP_DIV 7u
#define REBOP_OP_POW 8u
#define REBOP_OP_MAX 9u
#define REBOP_OP_MIN 10u
#define REBOP_OP_EXP 11u
#define REBOP_EXPR_MAX_DEPTH 32u
#define REBOP_EXPR_MAX_CODE 4096u

#define REBOP_EXPR_TOKEN(op, arg) ((uint32_t)(op) | ((uint32_t)(arg) << 8))
#define REBOP_EXPR_OPCODE(tok) ((tok)&0xffu)
#define REBOP_EXPR_ARG(tok) ((tok) >> 8)

/* 1/n!, n = 0..13 (Taylor coefficients of exp) */
#define REBOP_EXPC13 0x1.6124613a86d09p-33
#define REBOP_EXPC12 0x1.1eed8eff8d898p-29
#define REBOP_EXPC11 0x1.ae64567f544e4p-26
#define REBOP_EXPC10 0x1.27e4fb7789f5cp-22
#define REBOP_EXPC9 0x1.71de3a556c734p-19
#define REBOP_EXPC8 0x1.a01a01a01a01ap-16
#define REBOP_EXPC7 0x1.a01a01a01a01ap-13
#define REBOP_EXPC6 0x1.6c16c16c16c17p-10
#define REBOP_EXPC5 0x1.1111111111111p-7
#define REBOP_EXPC4 0x1.5555555555555p-5
#define REBOP_EXPC3 0x1.5555555555555p-3
#define REBOP_EXPC2 0x1.0000000000000p-1
#define REBOP_INV_LN2 0x1.71547652b82fep+0
/* exp(x) overflows above ln(DBL_MAX) and is below half the least subnormal
 * under ln(2^-1075) */
#define REBOP_EXP_OVER 0x1.62e42fefa39efp+9
#define REBOP_EXP_UNDER (-0x1.74910d52d3052p+9)
#define REBOP_ROUND_SHIFT 0x1.8p52

/* 2^e for e in [-1022, 1023] */
REBOP_HD double rebop_pow2i(int32_t e) {
  return rebop_f64_from_bits((uint64_t)(int64_t)(e + 1023) << 52);
}

REBOP_HD double rebop_exp64(double x) {
  if (x != x) return x;
  if (x > REBOP_EXP_OVER) return 1.0 / 0.0;
  if (x < REBOP_EXP_UNDER) return 0.0;
  /* k = nearest integer to x / ln 2 (round-to-nearest via the 1.5 * 2^52 shift) */
  const double kd = (x * REBOP_INV_LN2 + REBOP_ROUND_SHIFT) - REBOP_ROUND_SHIFT;
  const int32_t k = (int32_t)kd;
  const double r = (x - kd * REBOP_LN2_HI) - kd * REBOP_LN2_LO;
  double p = REBOP_EXPC13;
  p = __builtin_fma(p, r, REBOP_EXPC12);
  p = __builtin_fma(p, r, REBOP_EXPC11);
  p = __builtin_fma(p, r, REBOP_EXPC10);
  p = __builtin_fma(p, r, REBOP_EXPC9);
  p = __builtin_fma(p, r, REBOP_EXPC8);
  p = __builtin_fma(p, r, REBOP_EXPC7);
  p = __builtin_fma(p, r, REBOP_EXPC6);
  p = __builtin_fma(p, r, REBOP_EXPC5);
  p = __builtin_fma(p, r, REBOP_EXPC4);
  p = __builtin_fma(p, r, REBOP_EXPC3);
  p = __builtin_fma(p, r, REBOP_EXPC2);
  p = __builtin_fma(p, r, 1.0);
  p = __builtin_fma(p, r, 1.0);
  /* p * 2^k with k in [-1075, 1024]: two exact-range factors */
  const int32_t k1 = k / 2;
  const int32_t k2 = k - k1;
  return (p * rebop_pow2i(k1)) * rebop_pow2i(k2);
}

/* 1 if y is an integer (finite), else 0 */
REBOP_HD int rebop_is_integral(double y) {
  return __builtin_trunc(y) == y && y - y == 0.0;
}

/* for integral y: 1 if odd */
REBOP_HD int rebop_is_odd(double y) {
  const double a = __builtin_fabs(y);
  if (a >= 0x1p53) return 0;
  return ((uint64_t)a & 1u) != 0;
}

REBOP_HD double rebop_pow64(double x, double y) {
  if (y == 0.0) return 1.0;
  if (x == 1.0) return 1.0;
  if (x != x || y != y) return x + y;
  const double ay = __builtin_fabs(y);
  const double ax = __builtin_fabs(x);
  const double inf = 1.0 / 0.0;
  if (ay == inf) {
    if (ax == 1.0) return 1.0;
    if ((ax < 1.0) == (y > 0.0)) return 0.0;
    return inf;
  }
  const int yint = rebop_is_integral(y);
  const int yodd = yint && rebop_is_odd(y);
  if (x == 0.0) {
    if (y < 0.0) return yodd ? 1.0 / x : inf;
    return yodd ? x : 0.0;
  }
  if (ax == inf) {
    if (x > 0.0) return y < 0.0 ? 0.0 : inf;
    if (y < 0.0) return yodd ? -0.0 : 0.0;
    return yodd ? -inf : inf;
  }
  if (x < 0.0 && !yint) return (x - x) / (x - x); /* NaN */
  double mag;
  if (yint && ay <= 64.0) {
    uint32_t n = (uint32_t)ay;
    double b = ax, acc = 1.0;
    while (n) {
      if (n & 1u) acc = acc * b;
      n >>= 1;
      if (n) b = b * b;
    }
    mag = y < 0.0 ? 1.0 / acc : acc;
  } else {
    /* ln64 takes normal arguments: scale a subnormal |x| by 2^54 first */
    double lx;
    if (ax < 0x1p-1022)
      lx = rebop_ln64(ax * 0x1p54) - (54.0 * REBOP_LN2_HI + 54.0 * REBOP_LN2_LO);
    else
      lx = rebop_ln64(ax);
    mag = rebop_exp64(y * lx);
  }
  return (x < 0.0 && yodd) ? -mag : mag;
}

REBOP_HD double rebop_fmax64(double a, double b) {
  if (a != a) return b;
  if (b != b) return a;
  return a > b ? a : b;
}

REBOP_HD double rebop_fmin64(double a, double b) {
  if (a != a) return b;
  if (b != b) return a;
  return a < b ? a : b;
}

#endif /* REBOP_EXPR_H */
