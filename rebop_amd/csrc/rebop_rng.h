/*
 * rebop_rng.h -- the random stream of the batched direct-method SSA.
 *
 * One header, compiled by gcc (host, C11) and by hipcc (gfx950 device code),
 * so that host and device draw bit-identical numbers.  Everything here is
 * integer arithmetic or IEEE-754 binary64 operations that are correctly
 * rounded on both targets (+, -, *, /, fma); both builds use
 * -ffp-contract=off, and the clang pragma below repeats that for hipcc.
 *
 * What the reference draws, and what replaces it:
 *   reference (src/gillespie.rs:327)  t += rng.sample(Exp1) / total_rate
 *   reference (src/gillespie.rs:332)  chosen = total_rate * rng.random::<f64>()
 *   with rng = rand 0.10.2 SmallRng and rand_distr 0.6.0 Exp1 (ziggurat).
 * Here one Philox4x32-10 block per exponential draw supplies both numbers:
 *   key     = (seed lo32, seed hi32)
 *   counter = (draw lo32, draw hi32, traj lo32, traj hi32)
 *   E = -ln64(U1), U1 = ((w1:w0 >> 11) + 1) * 2^-53   in (0, 1]
 *   U = (w3:w2 >> 11) * 2^-53                           in [0, 1)
 * The draw index counts exponential draws of one trajectory, so an overshoot
 * (the reference consumes and discards the Exp1 draw, gillespie.rs:327-330)
 * advances the counter by one and leaves that block's U unused.
 *
 * Philox4x32-10: Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
 * as 1, 2, 3", SC'11.  Multipliers and Weyl increments are the published ones.
 *
 * ln64: argument reduction x = 2^k * m, m in [sqrt(1/2), sqrt(2)), f = m - 1
 * (exact by Sterbenz), s = f / (2 + f), ln(1 + f) = 2 atanh(s)
 *   = f - (f^2/2 - s (f^2/2 + z g(z))),  z = s^2,
 * where g(z) ~ (2 atanh(sqrt z)/sqrt z - 2) / z = 2/3 + 2z/5 + ... is a degree-7
 * relative-minimax polynomial on [0, 0.02944] (relative error 2^-58.2) whose
 * coefficients were fitted for this file with mpmath (Remez exchange), and
 * ln 2 = LN2_HI + LN2_LO with LN2_HI = ln 2 rounded to 32 significant bits.
 * (The f - (hfsq - s(hfsq + R)) arrangement keeps the exact term f outside all
 * rounding; it is the textbook atanh reduction, e.g. J.-M. Muller, "Elementary
 * Functions", ch. 11.)  Accuracy is checked against mpmath by the test suite.
 */
#ifndef REBOP_RNG_H
#define REBOP_RNG_H

#include <stdint.h>

#if defined(__HIPCC__)
#define REBOP_HD __host__ __device__ static inline __attribute__((always_inline))
#else
#define REBOP_HD static inline __attribute__((always_inline))
#endif

#if defined(__clang__)
#pragma clang fp contract(off)
#endif

/* ---------------- Philox4x32-10 ---------------- */

#define REBOP_PHILOX_M0 0xD2511F53u
#define REBOP_PHILOX_M1 0xCD9E8D57u
#define REBOP_PHILOX_W0 0x9E3779B9u
#define REBOP_PHILOX_W1 0xBB67AE85u

typedef struct {
  uint32_t w[4];
} rebop_u32x4;

/* a ^ b ^ c: one v_bitop3_b32 (truth table 0x96) in gfx950 device code */
#if defined(__HIP_DEVICE_COMPILE__) && defined(__gfx950__)
#define REBOP_XOR3(a, b, c) __builtin_amdgcn_bitop3_b32((a), (b), (c), 0x96)
#else
#define REBOP_XOR3(a, b, c) ((a) ^ (b) ^ (c))
#endif

/* One Philox4x32-10 block: counter (c0..c3), key (k0, k1). */
REBOP_HD rebop_u32x4 rebop_philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                         uint32_t k0, uint32_t k1) {
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
  for (int round = 0; round < 10; ++round) {
    const uint64_t p0 = (uint64_t)REBOP_PHILOX_M0 * (uint64_t)c0;
    const uint64_t p1 = (uint64_t)REBOP_PHILOX_M1 * (uint64_t)c2;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    const uint32_t n0 = REBOP_XOR3(hi1, c1, k0);
    const uint32_t n2 = REBOP_XOR3(hi0, c3, k1);
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
    k0 += REBOP_PHILOX_W0;
    k1 += REBOP_PHILOX_W1;
  }
  rebop_u32x4 r;
  r.w[0] = c0;
  r.w[1] = c1;
  r.w[2] = c2;
  r.w[3] = c3;
  return r;
}

/* The block of draw `draw` of trajectory `traj` under `seed`. */
REBOP_HD rebop_u32x4 rebop_draw_block(uint64_t seed, uint64_t traj, uint64_t draw) {
  return rebop_philox4x32_10((uint32_t)draw, (uint32_t)(draw >> 32), (uint32_t)traj,
                             (uint32_t)(traj >> 32), (uint32_t)seed, (uint32_t)(seed >> 32));
}

/* ---------------- uniform mappings ---------------- */

/* (hi:lo >> 11) is a 53-bit integer; its conversion to double is exact. */
REBOP_HD double rebop_u01_closed_open(uint32_t lo, uint32_t hi) {
  const uint64_t j = ((((uint64_t)hi) << 32) | (uint64_t)lo) >> 11;
  return (double)j * 0x1p-53;
}

REBOP_HD double rebop_u01_open_closed(uint32_t lo, uint32_t hi) {
  const uint64_t j = (((((uint64_t)hi) << 32) | (uint64_t)lo) >> 11) + 1u;
  return (double)j * 0x1p-53;
}

/* ---------------- ln64 ---------------- */

#define REBOP_LN2_HI 0x1.62e42ff000000p-1
#define REBOP_LN2_LO (-0x1.718432a1b0e26p-35)
#define REBOP_LNG0 0x1.5555555555555p-1
#define REBOP_LNG1 0x1.9999999999a38p-2
#define REBOP_LNG2 0x1.2492492476b57p-2
#define REBOP_LNG3 0x1.c71c72017efabp-3
#define REBOP_LNG4 0x1.745cf8f5e3405p-3
#define REBOP_LNG5 0x1.3b1c384439338p-3
#define REBOP_LNG6 0x1.0fbe4c71ea87dp-3
#define REBOP_LNG7 0x1.0c06728152d22p-3
/* mantissa field of sqrt(2) rounded to binary64 */
#define REBOP_SQRT2_MANT 0x6a09e667f3bcdull

REBOP_HD uint64_t rebop_f64_bits(double x) {
  uint64_t u;
  __builtin_memcpy(&u, &x, sizeof u);
  return u;
}

REBOP_HD double rebop_f64_from_bits(uint64_t u) {
  double x;
  __builtin_memcpy(&x, &u, sizeof x);
  return x;
}

/*
 * Natural logarithm for positive, normal, finite x (the SSA only calls it on
 * U in [2^-53, 1]).  No special-value handling: zero, subnormals, negative
 * numbers, inf and NaN are outside the contract.
 */
REBOP_HD double rebop_ln64(double x) {
  const uint64_t u = rebop_f64_bits(x);
  int32_t k = (int32_t)(u >> 52) - 1023;
  uint64_t mant = u & 0x000fffffffffffffull;
  uint64_t mexp = 0x3ff0000000000000ull; /* m in [1, 2) */
  if (mant >= REBOP_SQRT2_MANT) {       /* m in [sqrt2, 2): use m/2, k+1 */
    mexp = 0x3fe0000000000000ull;
    k += 1;
  }
  const double m = rebop_f64_from_bits(mexp | mant);
  const double f = m - 1.0;
  const double s = f / (2.0 + f);
  const double z = s * s;
  double g = REBOP_LNG7;
  g = __builtin_fma(g, z, REBOP_LNG6);
  g = __builtin_fma(g, z, REBOP_LNG5);
  g = __builtin_fma(g, z, REBOP_LNG4);
  g = __builtin_fma(g, z, REBOP_LNG3);
  g = __builtin_fma(g, z, REBOP_LNG2);
  g = __builtin_fma(g, z, REBOP_LNG1);
  g = __builtin_fma(g, z, REBOP_LNG0);
  const double R = z * g;
  const double hfsq = 0.5 * f * f;
  const double dk = (double)k;
  /* ln x = k ln2_hi + (f - (hfsq - (s (hfsq + R) + k ln2_lo))) */
  const double corr = s * (hfsq + R) + dk * REBOP_LN2_LO;
  return dk * REBOP_LN2_HI + (f - (hfsq - corr));
}

/* The exponential E of draw block b: E = -ln(U1), U1 in (0, 1]. */
REBOP_HD double rebop_exp1_of(rebop_u32x4 b) {
  return -rebop_ln64(rebop_u01_open_closed(b.w[0], b.w[1]));
}

/* The uniform U of draw block b, in [0, 1). */
REBOP_HD double rebop_unif_of(rebop_u32x4 b) {
  return rebop_u01_closed_open(b.w[2], b.w[3]);
}

#endif /* REBOP_RNG_H */
