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
 * ln64: table-driven reduction (P. T. P. Tang, "Table-driven implementation
 * of the logarithm function in IEEE floating-point arithmetic", ACM TOMS 16,
 * 1990): x = 2^k m with m in [0.70703125, 1.4140625); the top bits of m pick
 * one of 128 buckets with a stored invc ~ 1/c (c the bucket's midpoint, 1.0
 * for the two buckets touching 1.0) and -ln(invc) as a double pair; then
 * r = m invc - 1 (one fma, |r| < 2^-7) and
 *   ln x = k ln2 + (-ln invc) + r - r^2/2 + r^3 P(r),
 * P the Taylor polynomial of (log1p(r) - r + r^2/2) / r^3 to degree 5
 * (truncation < 2^-45 relative to r^3).  k ln2_hi + (-ln invc)_hi is summed
 * exactly (Fast2Sum) and its rounding error carried with the low parts.  No
 * division.  The table comes from gen_ln_table.py (mpmath); accuracy against
 * mpmath is checked by the test suite (faithful: < 1 ulp).
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

/* The draws of one trajectory: the counter is (draw_lo, draw_hi, traj_lo,
 * traj_hi), so the first round's product of the trajectory half, M1 * traj_lo,
 * and the key words it meets are the same for every draw.  They are taken
 * once per trajectory here; rebop_stream_block finishes a block from them
 * (a 32 x 32 -> 64 product and the trajectory id's addition fewer per draw,
 * which the one-lane-per-trajectory loops use; measured 1.4 % on the
 * 4-species dimer launch).  Same function as rebop_philox4x32_10, term for
 * term. */
typedef struct {
  uint32_t h0; /* hi(M1 * traj_lo) ^ k0: round 0's new c0, less draw_hi */
  uint32_t l1; /* lo(M1 * traj_lo): round 0's new c1 */
  uint32_t g2; /* traj_hi ^ k1: round 0's new c2, less hi(M0 * draw_lo) */
  uint32_t k0, k1;
} rebop_stream;

REBOP_HD rebop_stream rebop_stream_of(uint64_t seed, uint64_t traj) {
  rebop_stream s;
  const uint64_t p1 = (uint64_t)REBOP_PHILOX_M1 * (uint64_t)(uint32_t)traj;
  s.k0 = (uint32_t)seed;
  s.k1 = (uint32_t)(seed >> 32);
  s.h0 = (uint32_t)(p1 >> 32) ^ s.k0;
  s.l1 = (uint32_t)p1;
  s.g2 = (uint32_t)(traj >> 32) ^ s.k1;
  return s;
}

REBOP_HD rebop_u32x4 rebop_stream_block(rebop_stream s, uint64_t draw) {
  uint32_t k0 = s.k0, k1 = s.k1;
  const uint64_t p0 = (uint64_t)REBOP_PHILOX_M0 * (uint64_t)(uint32_t)draw;
  uint32_t c0 = s.h0 ^ (uint32_t)(draw >> 32);
  uint32_t c1 = s.l1;
  uint32_t c2 = (uint32_t)(p0 >> 32) ^ s.g2;
  uint32_t c3 = (uint32_t)p0;
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
  for (int round = 1; round < 10; ++round) {
    k0 += REBOP_PHILOX_W0;
    k1 += REBOP_PHILOX_W1;
    const uint64_t q0 = (uint64_t)REBOP_PHILOX_M0 * (uint64_t)c0;
    const uint64_t q1 = (uint64_t)REBOP_PHILOX_M1 * (uint64_t)c2;
    const uint32_t n0 = REBOP_XOR3((uint32_t)(q1 >> 32), c1, k0);
    const uint32_t n2 = REBOP_XOR3((uint32_t)(q0 >> 32), c3, k1);
    c0 = n0;
    c1 = (uint32_t)q1;
    c2 = n2;
    c3 = (uint32_t)q0;
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
/* bit pattern of 0.70703125, the lower end of the reduced range */
#define REBOP_LN_BASE 0x3FE6A00000000000ull
/* mantissa field above which m is halved (m >= 1.4140625) */
#define REBOP_LN_HALVE 0x6a00000000000ull

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

/* {invc, (-ln invc) hi, (-ln invc) lo, 0} per bucket (gen_ln_table.py) */
#if defined(__HIP_DEVICE_COMPILE__)
static __constant__ const double rebop_ln_table[128][4] = {
#else
static const double rebop_ln_table[128][4] = {
#endif
    {0x1.691473a88d0c0p+0, -0x1.602d08af091ecp-2, -0x1.a45db7cfd9230p-56, 0.0},
    {0x1.6719f3601671ap+0, -0x1.5a8cadbbedfa1p-2, -0x1.64f5081307f22p-60, 0.0},
    {0x1.6524f853b4aa3p+0, -0x1.54f431b7be1a8p-2, 0x1.0b3f6ef6ae452p-58, 0.0},
    {0x1.63356b88ac0dep+0, -0x1.4f637ebba9810p-2, 0x1.68cb3124b9245p-56, 0.0},
    {0x1.614b36831ae94p+0, -0x1.49da7f3bcc420p-2, 0x1.d964a168ccacbp-57, 0.0},
    {0x1.5f66434292dfcp+0, -0x1.44591e0539f49p-2, -0x1.a76d6dc2782dap-59, 0.0},
    {0x1.5d867c3ece2a5p+0, -0x1.3edf463c1683ep-2, 0x1.c852fe587def8p-57, 0.0},
    {0x1.5babcc647fa91p+0, -0x1.396ce359bbf53p-2, 0x1.5c5663663d163p-59, 0.0},
    {0x1.59d61f123ccaap+0, -0x1.3401e12aecba0p-2, -0x1.f95523adc5c9fp-57, 0.0},
    {0x1.5805601580560p+0, -0x1.2e9e2bce12286p-2, 0x1.f3ed72e23e134p-57, 0.0},
    {0x1.56397ba7c52e2p+0, -0x1.2941afb186b7cp-2, -0x1.6a4678ebaa300p-59, 0.0},
    {0x1.54725e6bb82fep+0, -0x1.23ec5991eba49p-2, -0x1.76eba35bbf0dfp-61, 0.0},
    {0x1.52aff56a8054bp+0, -0x1.1e9e1678899f5p-2, -0x1.64b0dd2687939p-58, 0.0},
    {0x1.50f22e111c4c5p+0, -0x1.1956d3b9bc2f9p-2, -0x1.0e75a3542856fp-58, 0.0},
    {0x1.4f38f62dd4c9bp+0, -0x1.14167ef367784p-2, -0x1.ef824daaf53e9p-56, 0.0},
    {0x1.4d843bedc2c4cp+0, -0x1.0edd060b78082p-2, -0x1.2d4b610d7d4f5p-57, 0.0},
    {0x1.4bd3edda68fe1p+0, -0x1.09aa572e6c6d4p-2, -0x1.f9e17343426a9p-56, 0.0},
    {0x1.4a27fad76014ap+0, -0x1.047e60cde83b7p-2, -0x1.08869cbf9e344p-56, 0.0},
    {0x1.4880522014880p+0, -0x1.feb2233ea07cbp-3, -0x1.8de00938b4c30p-61, 0.0},
    {0x1.46dce34596066p+0, -0x1.f474b134df228p-3, 0x1.9f1df7b5daab7p-60, 0.0},
    {0x1.453d9e2c776cap+0, -0x1.ea4449f04aaf5p-3, 0x1.f33919ab94074p-57, 0.0},
    {0x1.43a2730abee4dp+0, -0x1.e020cc6235ab5p-3, 0x1.f0adb91423f18p-57, 0.0},
    {0x1.420b5265e5951p+0, -0x1.d60a17f903514p-3, 0x1.50df841a71b7ap-57, 0.0},
    {0x1.40782d10e6566p+0, -0x1.cc000c9db3c52p-3, -0x1.67a2a8500729ep-58, 0.0},
    {0x1.3ee8f42a5af07p+0, -0x1.c2028ab17f9b5p-3, -0x1.c11aa3853a5f0p-57, 0.0},
    {0x1.3d5d991aa75c6p+0, -0x1.b811730b823d4p-3, 0x1.d7c46328983c6p-58, 0.0},
    {0x1.3bd60d9232955p+0, -0x1.ae2ca6f672bd8p-3, 0x1.a4a356155f779p-57, 0.0},
    {0x1.3a524387ac822p+0, -0x1.a454082e6ab03p-3, 0x1.e0df823a3cb3dp-58, 0.0},
    {0x1.38d22d366088ep+0, -0x1.9a8778debaa3ap-3, -0x1.28fbfb0e3f0fcp-58, 0.0},
    {0x1.3755bd1c945eep+0, -0x1.90c6db9fcbcdbp-3, 0x1.357718d7ca4cfp-58, 0.0},
    {0x1.35dce5f9f2af8p+0, -0x1.871213750e994p-3, 0x1.a97a0ca115d60p-57, 0.0},
    {0x1.34679ace01346p+0, -0x1.7d6903caf5acdp-3, 0x1.0b17c301d6e14p-57, 0.0},
    {0x1.32f5ced6a1dfap+0, -0x1.73cb9074fd14dp-3, 0x1.721a000b4cf01p-57, 0.0},
    {0x1.3187758e9ebb6p+0, -0x1.6a399dabbd383p-3, -0x1.76332bd4b341fp-57, 0.0},
    {0x1.301c82ac40260p+0, -0x1.60b3100b09474p-3, -0x1.526cee0fd7f4ap-57, 0.0},
    {0x1.2eb4ea1fed14bp+0, -0x1.5737cc9018cddp-3, 0x1.00b28ef013c72p-57, 0.0},
    {0x1.2d50a012d50a0p+0, -0x1.4dc7b897bc1c7p-3, -0x1.b60ae1ff0e82ep-59, 0.0},
    {0x1.2bef98e5a3711p+0, -0x1.4462b9dc9b3dcp-3, 0x1.85388d830c709p-59, 0.0},
    {0x1.2a91c92f3c105p+0, -0x1.3b08b6757f2a7p-3, -0x1.5e1ad9be0a4cdp-57, 0.0},
    {0x1.293725bb804a5p+0, -0x1.31b994d3a4f86p-3, 0x1.1238b5efe0665p-57, 0.0},
    {0x1.27dfa38a1ce4dp+0, -0x1.28753bc11aba2p-3, 0x1.7394d9fa33313p-57, 0.0},
    {0x1.268b37cd60127p+0, -0x1.1f3b925f25d44p-3, -0x1.08b27be4e6b15p-57, 0.0},
    {0x1.2539d7e9177b2p+0, -0x1.160c8024b27b0p-3, 0x1.355bfd870afebp-59, 0.0},
    {0x1.23eb79717605bp+0, -0x1.0ce7ecdccc28bp-3, -0x1.1b57fea88da98p-59, 0.0},
    {0x1.22a0122a0122ap+0, -0x1.03cdc0a51ec0dp-3, -0x1.19e2d3f8b7d10p-57, 0.0},
    {0x1.21579804855e6p+0, -0x1.f57bc7d9005dbp-4, 0x1.d361574fb24e2p-58, 0.0},
    {0x1.2012012012012p+0, -0x1.e3707ee30487bp-4, -0x1.9399d9aaf3b33p-59, 0.0},
    {0x1.1ecf43c7fb84cp+0, -0x1.d179788219362p-4, 0x1.b12841044a96cp-58, 0.0},
    {0x1.1d8f5672e4abdp+0, -0x1.bf968769fca18p-4, 0x1.06e4fb7af9c69p-58, 0.0},
    {0x1.1c522fc1ce059p+0, -0x1.adc77ee5aea8ep-4, -0x1.d7d8f39bee658p-58, 0.0},
    {0x1.1b17c67f2bae3p+0, -0x1.9c0c32d4d254dp-4, 0x1.627a0e199f569p-58, 0.0},
    {0x1.19e0119e0119ep+0, -0x1.8a6477a91dc29p-4, 0x1.3d4190a482421p-58, 0.0},
    {0x1.18ab083902bdbp+0, -0x1.78d02263d82d7p-4, -0x1.cbca5b4fdb87ep-58, 0.0},
    {0x1.1778a191bd684p+0, -0x1.674f089365a78p-4, -0x1.ca64e9980e048p-59, 0.0},
    {0x1.1648d50fc3201p+0, -0x1.55e10050e0382p-4, -0x1.9a0629e3973e4p-58, 0.0},
    {0x1.151b9a3fdd5c9p+0, -0x1.4485e03dbdfb0p-4, -0x1.3ba349aadbc6dp-58, 0.0},
    {0x1.13f0e8d344724p+0, -0x1.333d7f8183f4ap-4, 0x1.adaa06e211e9ep-59, 0.0},
    {0x1.12c8b89edc0acp+0, -0x1.2207b5c7854a1p-4, -0x1.b3f0431efb154p-58, 0.0},
    {0x1.11a3019a74826p+0, -0x1.10e45b3cae829p-4, -0x1.9b5ed72e6d974p-58, 0.0},
    {0x1.107fbbe011080p+0, -0x1.ffa6911ab9309p-5, 0x1.cd9f1f95c2ef1p-59, 0.0},
    {0x1.0f5edfab325a2p+0, -0x1.dda8adc67ee59p-5, 0x1.31936790bb3b2p-59, 0.0},
    {0x1.0e40655826011p+0, -0x1.bbcebfc68f424p-5, 0x1.cd1862f854848p-59, 0.0},
    {0x1.0d24456359e3ap+0, -0x1.9a187b573de81p-5, -0x1.b13b26f298a6ap-64, 0.0},
    {0x1.0c0a7868b4171p+0, -0x1.788595a3577c8p-5, -0x1.2f7c4c5b3c8bdp-62, 0.0},
    {0x1.0af2f722eecb5p+0, -0x1.5715c4c03cee1p-5, -0x1.5101dc4ebf91fp-59, 0.0},
    {0x1.09ddba6af8360p+0, -0x1.35c8bfaa13069p-5, 0x1.50830a65543a8p-63, 0.0},
    {0x1.08cabb37565e2p+0, -0x1.149e3e4005a8dp-5, 0x1.a9a4168fcebebp-60, 0.0},
    {0x1.07b9f29b8eae2p+0, -0x1.e72bf2813ce6ap-6, 0x1.8a4bba6a354fap-60, 0.0},
    {0x1.06ab59c7912fbp+0, -0x1.a55f548c5c427p-6, -0x1.f60d2fc36a0d9p-61, 0.0},
    {0x1.059eea0727586p+0, -0x1.63d6178690bbep-6, 0x1.18ed4d357c9dcp-60, 0.0},
    {0x1.04949cc1664c5p+0, -0x1.228fb1fea2e0ap-6, -0x1.3284991fe3d5cp-61, 0.0},
    {0x1.038c6b78247fcp+0, -0x1.c317384c75f0dp-7, -0x1.806208c04c21fp-61, 0.0},
    {0x1.02864fc7729e9p+0, -0x1.41929f968330cp-7, -0x1.3aae809b43dd0p-61, 0.0},
    {0x1.0182436517a37p+0, -0x1.8121214586b02p-8, 0x1.c7d68c0d910f2p-62, 0.0},
    {0x1.0000000000000p+0, 0x0.0p+0, 0x0.0p+0, 0.0},
    {0x1.0000000000000p+0, 0x0.0p+0, 0x0.0p+0, 0.0},
    {0x1.fa11caa01fa12p-1, 0x1.7dc475f810a69p-7, 0x1.74944bc161072p-61, 0.0},
    {0x1.f6310aca0dbb5p-1, 0x1.3cea44346a584p-6, -0x1.865ad48159d00p-61, 0.0},
    {0x1.f25f644230ab5p-1, 0x1.b9fc027af919ap-6, -0x1.90ae69229dc86p-60, 0.0},
    {0x1.ee9c7f8458e02p-1, 0x1.1b0d98923d97fp-5, -0x1.74d7444dd6241p-59, 0.0},
    {0x1.eae807aba01ebp-1, 0x1.58a5bafc8e4d3p-5, -0x1.cab8569c56e40p-64, 0.0},
    {0x1.e741aa59750e4p-1, 0x1.95c830ec8e3f2p-5, 0x1.eb41d00a417e9p-60, 0.0},
    {0x1.e3a9179dc1a73p-1, 0x1.d276b8adb0b56p-5, 0x1.078f14c95ff53p-59, 0.0},
    {0x1.e01e01e01e01ep-1, 0x1.075983598e471p-4, 0x1.006d2999e22dcp-58, 0.0},
    {0x1.dca01dca01dcap-1, 0x1.253f62f0a1417p-4, 0x1.1f6d34e01d981p-61, 0.0},
    {0x1.d92f2231e7f8ap-1, 0x1.42edcbea646eep-4, -0x1.511583653349bp-58, 0.0},
    {0x1.d5cac807572b2p-1, 0x1.60658a93750c4p-4, -0x1.f108b1d8436d3p-59, 0.0},
    {0x1.d272ca3fc5b1ap-1, 0x1.7da766d7b12d0p-4, 0x1.a2240644d7da2p-59, 0.0},
    {0x1.cf26e5c44bfc6p-1, 0x1.9ab42462033aep-4, -0x1.a099e1c184e8ep-59, 0.0},
    {0x1.cbe6d9601cbe7p-1, 0x1.b78c82bb0eda0p-4, -0x1.3ef0e61f9b03cp-58, 0.0},
    {0x1.c8b265afb8a42p-1, 0x1.d4313d66cb35dp-4, 0x1.b90dd951d90fap-58, 0.0},
    {0x1.c5894d10d4986p-1, 0x1.f0a30c01162a4p-4, 0x1.8be64b8b7759bp-59, 0.0},
    {0x1.c26b5392ea01cp-1, 0x1.0671512ca596fp-3, -0x1.2f39b81479b67p-58, 0.0},
    {0x1.bf583ee868d8bp-1, 0x1.14785846742acp-3, 0x1.94409f1d3f83ap-60, 0.0},
    {0x1.bc4fd65883e7bp-1, 0x1.2266f190a5acdp-3, -0x1.dab840e7f6177p-57, 0.0},
    {0x1.b951e2b18ff23p-1, 0x1.303d718e47fd5p-3, -0x1.b5ae71f658247p-57, 0.0},
    {0x1.b65e2e3beee05p-1, 0x1.3dfc2b0ecc62ap-3, 0x1.ba62b8c13f7f4p-57, 0.0},
    {0x1.b37484ad806cep-1, 0x1.4ba36f39a55e5p-3, -0x1.f767e433c98aap-57, 0.0},
    {0x1.b094b31d922a4p-1, 0x1.59338d9982085p-3, 0x1.8d16eaaba9419p-57, 0.0},
    {0x1.adbe87f94905ep-1, 0x1.66acd4272ad51p-3, -0x1.9201c9c3d5165p-59, 0.0},
    {0x1.aaf1d2f87ebfdp-1, 0x1.740f8f54037a3p-3, 0x1.6d9bf9d57b326p-58, 0.0},
    {0x1.a82e65130e159p-1, 0x1.815c0a14357e9p-3, 0x1.141b7f8c5fa9ep-58, 0.0},
    {0x1.a574107688a4ap-1, 0x1.8e928de886d41p-3, 0x1.2589eb96a6240p-59, 0.0},
    {0x1.a2c2a87c51ca0p-1, 0x1.9bb362e7dfb85p-3, -0x1.51439c1ff83e7p-58, 0.0},
    {0x1.a01a01a01a01ap-1, 0x1.a8becfc882f19p-3, -0x1.a8c37918c39ebp-58, 0.0},
    {0x1.9d79f176b682dp-1, 0x1.b5b519e8fb5a6p-3, -0x1.d5d8023e61e5fp-57, 0.0},
    {0x1.9ae24ea5510dap-1, 0x1.c2968558c18c2p-3, 0x1.6108e3ae024acp-60, 0.0},
    {0x1.9852f0d8ec0ffp-1, 0x1.cf6354e09c5ddp-3, 0x1.339a07d55b696p-57, 0.0},
    {0x1.95cbb0be377aep-1, 0x1.dc1bca0abec7bp-3, 0x1.c698a33316dfbp-58, 0.0},
    {0x1.934c67f9b2ce6p-1, 0x1.e8c0252aa5a60p-3, -0x1.dc074737f9135p-60, 0.0},
    {0x1.90d4f120190d5p-1, 0x1.f550a564b7b37p-3, -0x1.13a09202fe73dp-57, 0.0},
    {0x1.8e6527af1373fp-1, 0x1.00e6c45ad501dp-2, -0x1.3b9568ff6feadp-57, 0.0},
    {0x1.8bfce8062ff3ap-1, 0x1.071b85fcd590dp-2, 0x1.08b83fcbdef40p-57, 0.0},
    {0x1.899c0f601899cp-1, 0x1.0d46b579ab74bp-2, 0x1.21f640e1e5ec9p-56, 0.0},
    {0x1.87427bcc092b9p-1, 0x1.136870293a8b0p-2, 0x1.86cc531dba494p-57, 0.0},
    {0x1.84f00c2780614p-1, 0x1.1980d2dd4236fp-2, -0x1.02c2e4f1b2eb9p-56, 0.0},
    {0x1.82a4a0182a4a0p-1, 0x1.1f8ff9e48a2f3p-2, -0x1.93fbf3418960dp-57, 0.0},
    {0x1.8060180601806p-1, 0x1.2596010df763ap-2, -0x1.9eed8ae0ebd3cp-59, 0.0},
    {0x1.7e225515a4f1dp-1, 0x1.2b9303ab89d25p-2, -0x1.85ad7f614ab51p-58, 0.0},
    {0x1.7beb3922e017cp-1, 0x1.31871c9544185p-2, -0x1.ea3598981366fp-57, 0.0},
    {0x1.79baa6bb6398bp-1, 0x1.3772662bfd85cp-2, 0x1.02a7589fba088p-57, 0.0},
    {0x1.77908119ac60dp-1, 0x1.3d54fa5c1f710p-2, 0x1.53668e578d9cdp-58, 0.0},
    {0x1.756cac201756dp-1, 0x1.432ef2a04e813p-2, -0x1.83262e2b59206p-57, 0.0},
    {0x1.734f0c541fe8dp-1, 0x1.49006804009d0p-2, -0x1.bff0d07c5df6dp-59, 0.0},
    {0x1.713786d9c7c09p-1, 0x1.4ec9732600269p-2, -0x1.1aa87d977dc5ep-56, 0.0},
    {0x1.6f26016f26017p-1, 0x1.548a2c3add263p-2, -0x1.58ce7bf1846eep-56, 0.0},
    {0x1.6d1a62681c861p-1, 0x1.5a42ab0f4cfe2p-2, -0x1.c6bcb7dee9a3dp-56, 0.0},
    {0x1.6b1490aa31a3dp-1, 0x1.5ff3070a793d4p-2, -0x1.063077d7e37b7p-56, 0.0},
};

/*
 * Natural logarithm for positive, normal, finite x (the SSA only calls it on
 * U in [2^-53, 1]).  No special-value handling: zero, subnormals, negative
 * numbers, inf and NaN are outside the contract.  `tab` is rebop_ln_table or
 * a copy of it (device code may pass one held in LDS).
 *
 * In two halves, so that a caller can issue the table read long before it
 * needs the value: rebop_ln64_begin reduces x and reads its table row,
 * rebop_ln64_finish evaluates the rest.  rebop_ln64_tab is the two in a row.
 */
typedef struct {
  double m, dk;       /* x = 2^k m */
  double t0, t1, t2;  /* the bucket's table row */
} rebop_ln_pre;

REBOP_HD rebop_ln_pre rebop_ln64_begin(double x, const double (*tab)[4]) {
  const uint64_t u = rebop_f64_bits(x);
  int32_t k = (int32_t)(u >> 52) - 1023;
  const uint64_t mant = u & 0x000fffffffffffffull;
  uint64_t mexp = 0x3ff0000000000000ull;
  if (mant >= REBOP_LN_HALVE) { /* m in [1.4140625, 2): use m/2, k+1 */
    mexp = 0x3fe0000000000000ull;
    k += 1;
  }
  const uint64_t mb = mexp | mant;
  const double *t = tab[(uint32_t)((mb - REBOP_LN_BASE) >> 45)];
  rebop_ln_pre p;
  p.m = rebop_f64_from_bits(mb);
  p.dk = (double)k;
  p.t0 = t[0];
  p.t1 = t[1];
  p.t2 = t[2];
  return p;
}

REBOP_HD double rebop_ln64_finish(rebop_ln_pre q) {
  const double r = __builtin_fma(q.m, q.t0, -1.0);
  double p = -0x1p-3;                            /* -1/8 */
  p = __builtin_fma(p, r, 0x1.2492492492492p-3);  /*  1/7 */
  p = __builtin_fma(p, r, -0x1.5555555555555p-3); /* -1/6 */
  p = __builtin_fma(p, r, 0x1.999999999999ap-3);  /*  1/5 */
  p = __builtin_fma(p, r, -0x1p-2);               /* -1/4 */
  p = __builtin_fma(p, r, 0x1.5555555555555p-2);  /*  1/3 */
  const double r2 = r * r;
  const double dk = q.dk;
  const double a = dk * REBOP_LN2_HI; /* exact: 32-bit ln2_hi times |k| <= 1075 */
  const double hi = a + q.t1;         /* |a| >= |t1| or a == 0: Fast2Sum */
  const double err = (a - hi) + q.t1;
  const double lo = __builtin_fma(dk, REBOP_LN2_LO, q.t2) + err;
  const double tail = __builtin_fma(r2, __builtin_fma(r, p, -0.5), lo);
  return hi + (r + tail);
}

REBOP_HD double rebop_ln64_tab(double x, const double (*tab)[4]) {
  return rebop_ln64_finish(rebop_ln64_begin(x, tab));
}

REBOP_HD double rebop_ln64(double x) { return rebop_ln64_tab(x, rebop_ln_table); }

/* The exponential E of draw block b: E = -ln(U1), U1 in (0, 1]. */
REBOP_HD double rebop_exp1_of(rebop_u32x4 b) {
  return -rebop_ln64(rebop_u01_open_closed(b.w[0], b.w[1]));
}

/* The same E with the ln table at `tab` (a copy of rebop_ln_table). */
REBOP_HD double rebop_exp1_of_tab(rebop_u32x4 b, const double (*tab)[4]) {
  return -rebop_ln64_tab(rebop_u01_open_closed(b.w[0], b.w[1]), tab);
}

/* The uniform U of draw block b, in [0, 1). */
REBOP_HD double rebop_unif_of(rebop_u32x4 b) {
  return rebop_u01_closed_open(b.w[2], b.w[3]);
}

#endif /* REBOP_RNG_H */
