// surrogate.hip -- kernels of the learned surrogate of the SSA.
//
// A multilayer perceptron maps (log rate constants, initial counts, t) of a
// Michaelis-Menten network to a K-component mixture per species (negative
// binomials, and binomials on the species' bound).  Its dense layers run on
// the matrix cores: BF16 operands, FP32
// accumulation (v_mfma_f32_16x16x32_bf16).  The loss head (mixture negative
// log-likelihood of the SSA samples and its gradient) runs in FP64.  Adam
// keeps FP32 master weights and writes the BF16 copy the GEMMs read.
//
// MFMA operand layout used below (16x16x32, wave64): lane l supplies
//   A[row = l % 16][k = 8 (l / 16) .. +8]   (8 bf16, k contiguous)
//   B[k = 8 (l / 16) .. +8][col = l % 16]
// and receives D[row = 4 (l / 16) + q][col = l % 16], q = 0..3.
// The GEMM test checks the product against a CPU reference.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include "../../include/rebop_surrogate.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

namespace sg {

// ---------------------------------------------------------------------------
// C = op(A) op(B), M x N x K, FP32 out.  !TA: A[m][k] (lda); TA: A stored
// [k][m].  !TB: B[k][n] (ldb); TB: B stored [n][k].  Tiles TM x TN x 32 (TM,
// TN = 64 or 128), four waves in 2 x 2, each (TM/2) x (TN/2) of 16 x 16 MFMA
// tiles: a 128 x 128 tile issues 16 MFMAs per wave per 8 fragment loads.
// M, N multiples of the tile, K of 32.  With gridDim.z > 1 (split K), block z
// covers k in [z*K/gz, (z+1)*K/gz) and adds its partial product into C
// (zeroed beforehand) atomically.  FUSED (gridDim.z == 1): instead of C, the
// epilogue writes act(C + bias) as BF16 to epi.H and/or FP32 to epi.F -- the
// arithmetic of bias_act_kernel on the same FP32 values, so the same bits,
// without the FP32 product's round trip through memory.
struct Epilogue {
  const float *bias;
  int relu;
  __bf16 *H;
  int ldh;
  float *F;
  int ldf;
};

template <bool TA, bool TB, int TM, int TN, bool FUSED>
__global__ void __launch_bounds__(256) gemm_kernel(int M, int N, int K, const __bf16 *__restrict__ A,
                                                   int lda, const __bf16 *__restrict__ B, int ldb,
                                                   float *__restrict__ C, int ldc, Epilogue epi) {
  constexpr int FM = TM / 32, FN = TN / 32; // 16-row fragments per wave
  __shared__ __attribute__((aligned(16))) __bf16 As[TM][40]; // [m][k], rows padded to 80 B
  __shared__ __attribute__((aligned(16))) __bf16 Bs[TN][40]; // [n][k]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
  const int wm = (wave >> 1) * (TM / 2), wn = (wave & 1) * (TN / 2);
  floatx4 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

  const int ksplit = K / (int)gridDim.z; // a multiple of 32 (checked by the host)
  const int kbeg = (int)blockIdx.z * ksplit, kend = kbeg + ksplit;
  for (int k0 = kbeg; k0 < kend; k0 += 32) {
#pragma unroll
    for (int p = 0; p < TM / 64; ++p) {
      const int idx = tid + 256 * p;
      if (!TA) {
        const int r = idx >> 2, c = (idx & 3) * 8;
        *(bf16x8 *)&As[r][c] = *(const bf16x8 *)&A[(size_t)(m0 + r) * lda + k0 + c];
      } else {
        const int kk = idx / (TM / 8), c = (idx % (TM / 8)) * 8;
        const bf16x8 v = *(const bf16x8 *)&A[(size_t)(k0 + kk) * lda + m0 + c];
#pragma unroll
        for (int q = 0; q < 8; ++q) As[c + q][kk] = v[q];
      }
    }
#pragma unroll
    for (int p = 0; p < TN / 64; ++p) {
      const int idx = tid + 256 * p;
      if (TB) {
        const int r = idx >> 2, c = (idx & 3) * 8;
        *(bf16x8 *)&Bs[r][c] = *(const bf16x8 *)&B[(size_t)(n0 + r) * ldb + k0 + c];
      } else {
        const int kk = idx / (TN / 8), c = (idx % (TN / 8)) * 8;
        const bf16x8 v = *(const bf16x8 *)&B[(size_t)(k0 + kk) * ldb + n0 + c];
#pragma unroll
        for (int q = 0; q < 8; ++q) Bs[c + q][kk] = v[q];
      }
    }
    __syncthreads();
    bf16x8 a[FM], b[FN];
#pragma unroll
    for (int i = 0; i < FM; ++i) a[i] = *(const bf16x8 *)&As[wm + i * 16 + (lane & 15)][(lane >> 4) * 8];
#pragma unroll
    for (int j = 0; j < FN; ++j) b[j] = *(const bf16x8 *)&Bs[wn + j * 16 + (lane & 15)][(lane >> 4) * 8];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int row = m0 + wm + i * 16 + (lane >> 4) * 4 + q;
        const int col = n0 + wn + j * 16 + (lane & 15);
        if (FUSED) {
          float v = acc[i][j][q] + epi.bias[col];
          if (epi.relu) v = v > 0.f ? v : 0.f;
          if (epi.H) epi.H[(size_t)row * epi.ldh + col] = (__bf16)v;
          if (epi.F) epi.F[(size_t)row * epi.ldf + col] = v;
        } else if (gridDim.z == 1) {
          C[(size_t)row * ldc + col] = acc[i][j][q];
        } else {
          atomicAdd(&C[(size_t)row * ldc + col], acc[i][j][q]);
        }
      }
}

// The same product for A [M][K] and B stored [N][K] (both k-contiguous, as
// the forward pass and the backward's dX use them), software-pipelined: the
// next 32-wide k slice is loaded from global memory into registers while the
// matrix cores work on the current one, which sits in one of two LDS
// buffers, so one barrier per slice separates the stores into the other
// buffer from its reads.  Same tiles, fragments and summation order per
// output as gemm_kernel<false, true>.
template <int TM, int TN, bool FUSED>
__global__ void __launch_bounds__(256) gemm_nt_kernel(int M, int N, int K, const __bf16 *__restrict__ A,
                                                      int lda, const __bf16 *__restrict__ B, int ldb,
                                                      float *__restrict__ C, int ldc, Epilogue epi) {
  constexpr int FM = TM / 32, FN = TN / 32, PA = TM / 64, PB = TN / 64;
  __shared__ __attribute__((aligned(16))) __bf16 As[2][TM][40];
  __shared__ __attribute__((aligned(16))) __bf16 Bs[2][TN][40];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
  const int wm = (wave >> 1) * (TM / 2), wn = (wave & 1) * (TN / 2);
  floatx4 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

  const int ksplit = K / (int)gridDim.z;
  const int kbeg = (int)blockIdx.z * ksplit, kend = kbeg + ksplit;
  // thread tid moves the 8-element chunks tid + 256 p: row (chunk >> 2), k (chunk & 3) * 8
  bf16x8 ra[PA], rb[PB];
  auto gload = [&](int k0) {
#pragma unroll
    for (int p = 0; p < PA; ++p) {
      const int idx = tid + 256 * p, r = idx >> 2, c = (idx & 3) * 8;
      ra[p] = *(const bf16x8 *)&A[(size_t)(m0 + r) * lda + k0 + c];
    }
#pragma unroll
    for (int p = 0; p < PB; ++p) {
      const int idx = tid + 256 * p, r = idx >> 2, c = (idx & 3) * 8;
      rb[p] = *(const bf16x8 *)&B[(size_t)(n0 + r) * ldb + k0 + c];
    }
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int p = 0; p < PA; ++p) {
      const int idx = tid + 256 * p, r = idx >> 2, c = (idx & 3) * 8;
      *(bf16x8 *)&As[buf][r][c] = ra[p];
    }
#pragma unroll
    for (int p = 0; p < PB; ++p) {
      const int idx = tid + 256 * p, r = idx >> 2, c = (idx & 3) * 8;
      *(bf16x8 *)&Bs[buf][r][c] = rb[p];
    }
  };
  gload(kbeg);
  lstore(0);
  __syncthreads();
  int buf = 0;
  for (int k0 = kbeg; k0 < kend; k0 += 32) {
    const bool more = k0 + 32 < kend;
    if (more) gload(k0 + 32); // in flight during the MFMAs below
    bf16x8 a[FM], b[FN];
#pragma unroll
    for (int i = 0; i < FM; ++i) a[i] = *(const bf16x8 *)&As[buf][wm + i * 16 + (lane & 15)][(lane >> 4) * 8];
#pragma unroll
    for (int j = 0; j < FN; ++j) b[j] = *(const bf16x8 *)&Bs[buf][wn + j * 16 + (lane & 15)][(lane >> 4) * 8];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    if (more) lstore(buf ^ 1); // that buffer's reads ended at the last barrier
    __syncthreads();
    buf ^= 1;
  }
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int row = m0 + wm + i * 16 + (lane >> 4) * 4 + q;
        const int col = n0 + wn + j * 16 + (lane & 15);
        if (FUSED) {
          float v = acc[i][j][q] + epi.bias[col];
          if (epi.relu) v = v > 0.f ? v : 0.f;
          if (epi.H) epi.H[(size_t)row * epi.ldh + col] = (__bf16)v;
          if (epi.F) epi.F[(size_t)row * epi.ldf + col] = v;
        } else if (gridDim.z == 1) {
          C[(size_t)row * ldc + col] = acc[i][j][q];
        } else {
          atomicAdd(&C[(size_t)row * ldc + col], acc[i][j][q]);
        }
      }
}

__global__ void zero_kernel(int M, int N, float *__restrict__ C, int ldc) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < (size_t)M * N) C[(i / N) * ldc + i % N] = 0.f;
}

// H = bf16(act(C + bias)), act = ReLU or identity; optionally the FP32 result too.
__global__ void bias_act_kernel(int M, int N, const float *__restrict__ C, int ldc,
                                const float *__restrict__ bias, int relu, __bf16 *__restrict__ H,
                                int ldh, float *__restrict__ F, int ldf) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)M * N) return;
  const int m = (int)(idx / N), n = (int)(idx % N);
  float v = C[(size_t)m * ldc + n] + bias[n];
  if (relu) v = v > 0.f ? v : 0.f;
  if (H) H[(size_t)m * ldh + n] = (__bf16)v;
  if (F) F[(size_t)m * ldf + n] = v;
}

// dZ = dH * (H > 0): the ReLU derivative from the stored activation; FP32 in
// place (for the bias gradient) and BF16 (for the next GEMMs).
__global__ void relu_grad_kernel(int M, int N, float *__restrict__ dH, int ldd,
                                 const __bf16 *__restrict__ H, int ldh, __bf16 *__restrict__ dZ,
                                 int ldz) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)M * N) return;
  const int m = (int)(idx / N), n = (int)(idx % N);
  const float h = (float)H[(size_t)m * ldh + n];
  const float g = h > 0.f ? dH[(size_t)m * ldd + n] : 0.f;
  dH[(size_t)m * ldd + n] = g;
  dZ[(size_t)m * ldz + n] = (__bf16)g;
}

// y[c][r] = bf16(x[r][c]) for an R x C matrix
__global__ void transpose_bf16_kernel(int R, int C, const float *__restrict__ x, __bf16 *__restrict__ y) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)R * C) return;
  const int r = (int)(i / C), c = (int)(i % C);
  y[(size_t)c * R + r] = (__bf16)x[i];
}

// Wp[k / 32][n][k % 32] = bf16(W[k][n]) for W [K][N] (K a multiple of 32)
__global__ void pack_kslices_kernel(int K, int N, const float *__restrict__ W, __bf16 *__restrict__ Wp) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)K * N) return;
  const int kk = (int)(i % 32), n = (int)((i / 32) % N), k = (int)(i / (32 * (size_t)N)) * 32 + kk;
  Wp[i] = (__bf16)W[(size_t)k * N + n];
}

__global__ void cast_bf16_kernel(size_t n, const float *__restrict__ x, __bf16 *__restrict__ y) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (__bf16)x[i];
}

// out[n] = sum_m X[m][n]: one thread per column, rows strided (coalesced).
// Column sums over a 2-D grid: block (x, y) sums columns 64x .. 64x+63 over
// rows 256y .. 256y+255 (four row groups of 64 in parallel, combined in LDS)
// and adds its partial sums into out (zeroed first) atomically.
__global__ void __launch_bounds__(256) colsum_kernel(int M, int N, const float *__restrict__ X,
                                                     int ldx, float *__restrict__ out) {
  __shared__ float part[4][64];
  const int c = threadIdx.x & 63, g = threadIdx.x >> 6;
  const int n = blockIdx.x * 64 + c;
  const int m0 = blockIdx.y * 256 + g * 64;
  float s = 0.f;
  if (n < N)
    for (int m = m0; m < m0 + 64 && m < M; ++m) s += X[(size_t)m * ldx + n];
  part[g][c] = s;
  __syncthreads();
  if (g == 0 && n < N) atomicAdd(&out[n], part[0][c] + part[1][c] + part[2][c] + part[3][c]);
}

__global__ void fill_kernel(int n, float *__restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = 0.f;
}

// Adam with FP32 master weights; writes the BF16 copy read by the GEMMs.
__global__ void adam_kernel(size_t n, float *__restrict__ w, const float *__restrict__ g,
                            float *__restrict__ m, float *__restrict__ v, __bf16 *__restrict__ wb,
                            float lr, float b1, float b2, float eps, float bc1, float bc2) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
  w[i] = wi;
  if (wb) wb[i] = (__bf16)wi;
}

// Inputs of item b: (point p, sample i) -> the REBOP_SG_FEATURES features,
// scaled to about [-1, 1]: log10 of the rate constants, E0, S0, t, their
// logs, and the logs of the kinetics' dimensionless groups (kcat t, kf E0 t,
// kf S0 t, K_M / S0 with K_M = (kb + kcat) / kf, kb / kcat, E0 / S0,
// kcat E0 t / S0); e0, s0 = the counts that bound the species (E0 for E and
// ES, S0 for S and P; exact in FP32).
// The item's inputs: rate constants, the bounding counts and the sample time.
struct ItemInputs {
  double kf, kb, kc, e0, s0, t;
};

__device__ __forceinline__ ItemInputs item_inputs(uint32_t p, uint32_t i, const double *__restrict__ rates,
                                                  const int64_t *__restrict__ x0, double tmax, int nb_steps) {
#pragma clang fp contract(off)
  ItemInputs v;
  v.kf = rates[3 * p + 0];
  v.kb = rates[3 * p + 1];
  v.kc = rates[3 * p + 2];
  v.e0 = (double)x0[4 * p + 0];
  v.s0 = (double)x0[4 * p + 1];
  v.t = tmax * (double)i / (double)nb_steps;
  return v;
}

// Log j of the seven the features use: log10 of kf, kb, kc, E0, S0, t and
// kb + kc.
__device__ __forceinline__ double item_log(const ItemInputs &v, int j) {
#pragma clang fp contract(off)
  switch (j) {
  case 0: return log10(v.kf);
  case 1: return log10(v.kb);
  case 2: return log10(v.kc);
  case 3: return log10(v.e0);
  case 4: return log10(v.s0);
  case 5: return log10(v.t);
  default: return log10(v.kb + v.kc);
  }
}

// Feature j (before the FP32 cast) from the inputs and their logs lg[0..6].
__device__ __forceinline__ double item_feature(const ItemInputs &v, const double *lg, double tmax, int j) {
#pragma clang fp contract(off)
  const double lf = lg[0], lb = lg[1], lc = lg[2], le = lg[3], ls = lg[4], lt = lg[5];
  switch (j) {
  case 0: return (lf + 1.0) * 0.5;
  case 1: return (lb + 1.0) * 0.5;
  case 2: return (lc + 1.0) * 0.5;
  case 3: return (v.e0 - 105.0) / 95.0;
  case 4: return (v.s0 - 525.0) / 475.0;
  case 5: return 2.0 * v.t / tmax - 1.0;
  case 6: return (le - 1.65) / 0.65;
  case 7: return (ls - 2.35) / 0.65;
  case 8: return (lt - 0.95) / 0.75;
  case 9: return (lc + lt) / 3.0;
  case 10: return (lf + le + lt - 1.6) / 3.4;
  case 11: return (lf + ls + lt - 2.3) / 3.4;
  case 12: return ((lg[6] - lf) - ls + 2.05) / 4.65; // K_M = (kb + kcat) / kf
  case 13: return (lb - lc) / 4.0;
  case 14: return (le - ls + 0.7) / 1.3;
  default: return (lc + le + lt - ls + 0.75) / 4.05;
  }
}

__device__ __forceinline__ void item_features(uint32_t p, uint32_t i, const double *__restrict__ rates,
                                              const int64_t *__restrict__ x0, double tmax, int nb_steps,
                                              float (&f)[REBOP_SG_FEATURES], float &e0f, float &s0f) {
  const ItemInputs v = item_inputs(p, i, rates, x0, tmax, nb_steps);
  double lg[7];
#pragma unroll
  for (int j = 0; j < 7; ++j) lg[j] = item_log(v, j);
#pragma unroll
  for (int j = 0; j < REBOP_SG_FEATURES; ++j) f[j] = (float)item_feature(v, lg, tmax, j);
  e0f = (float)v.e0;
  s0f = (float)v.s0;
}

// X[b][0..ldx) = item b's features, then zeros; bound[b][s] (nullable) = the
// count that bounds species s (E = E0 - ES, S <= S0, ES <= E0, P <= S0).
__global__ void features_kernel(int B, const uint32_t *__restrict__ item_point,
                                const uint32_t *__restrict__ item_sample, const double *__restrict__ rates,
                                const int64_t *__restrict__ x0, double tmax, int nb_steps,
                                __bf16 *__restrict__ X, int ldx, float *__restrict__ bound) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  float f[REBOP_SG_FEATURES], e0, s0;
  item_features(item_point[b], item_sample[b], rates, x0, tmax, nb_steps, f, e0, s0);
  for (int c = 0; c < ldx; ++c) X[(size_t)b * ldx + c] = (__bf16)(c < REBOP_SG_FEATURES ? f[c] : 0.f);
  if (bound) {
    bound[4 * b + 0] = e0;
    bound[4 * b + 1] = s0;
    bound[4 * b + 2] = e0;
    bound[4 * b + 3] = s0;
  }
}

// head[b][s*3K + K + c] += ln bound[b][s] for the negative-binomial
// components c < K/2: their log means (the binomial components' logits stay).
__global__ void head_offsets_kernel(int B, int S, int K, float *__restrict__ head, int ldh,
                                    const float *__restrict__ bound) {
  const int half = K / 2;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= B * S * half) return;
  const int b = idx / (S * half), s = (idx / half) % S, c = idx % half;
  head[(size_t)b * ldh + (size_t)s * 3 * K + K + c] += logf(bound[(size_t)b * S + s]);
}

// ---------------------------------------------------------------------------
// Inference through the whole network in one launch.  A workgroup of eight
// waves takes 128 items, whose activations stay in LDS (128 rows of up to 512
// BF16, stored in MFMA fragment order: see hs_offset) from the input layer
// to the last hidden layer: HBM sees the inputs and the head only.  Wave w computes
// all 128 rows and columns (N / 8) w .. +N/8 of each layer, so each weight
// fragment is fetched once per workgroup.  The weights come straight from
// global memory (every workgroup reads the same ones, so they stay in L2) in
// the k-slice-major layout Wp[k / 32][n][k % 32] (rebop_sg_pack_kslices_bf16):
// a wave's fragment of one 32-wide slice is 1 KB of consecutive bytes.  The
// MFMA takes the weights as A and the activations as B, so a lane's four
// outputs are four consecutive columns of one row (one 8-byte LDS store per
// fragment).  Per output, the arithmetic is gemm_nt_kernel's: the same
// 16x16x32 MFMAs over k slices 0, 32, 64, ... (a product of two BF16 values
// is exact in FP32, so which operand is A does not matter), then + bias,
// ReLU and the BF16 cast -- the bits of the layer-by-layer forward pass.
// FEAT: the input layer is item_features of items m0.. (columns >=
// REBOP_SG_FEATURES zero) and the last layer adds head_offsets_kernel's
// ln bound terms, so one launch turns (point, sample) items into mixture
// parameters; otherwise X [M][ldx] is the input.
constexpr int MLP_BM = 128, MLP_MAXW = 512, MLP_MAXL = 8;
constexpr int MLP_HEAD_LD = 128 * 4 + 16; // bytes per row of a staged 128-column head

// Byte offset of activation (m, n) in LDS.  Fragment order: the 16 x 32
// block of rows 16 (m / 16).. and columns 32 (n / 32).. is 1 KB, in which
// lane l's B operand (row l % 16, columns 8 (l / 16) .. +8 of the block)
// is 16 consecutive bytes at 16 l -- so a wave's fragment read is 1 KB of
// consecutive bytes, free of bank conflicts.  Row blocks are 16 KB apart
// (512 columns).
__device__ __forceinline__ uint32_t hs_offset(int m, int n) {
  return (uint32_t)(((m >> 4) * (MLP_MAXW / 32) + (n >> 5)) * 1024 + ((m & 15) + 16 * ((n & 31) >> 3)) * 16 +
                    (n & 7) * 2);
}

struct MlpNet {
  const __bf16 *Wp[MLP_MAXL]; // layer l: [dims[l] / 32][dims[l+1]][32]
  const float *bias[MLP_MAXL];
  int dims[MLP_MAXL + 1]; // dims[0] multiple of 64 up to 512, the others 128, 256 or 512
  int depth;
};

struct MlpItems {
  const uint32_t *item_point, *item_sample;
  const double *rates;
  const int64_t *x0;
  double tmax;
  int nb_steps;
};

// Layer l's table entries, selected without indexing the argument block (a
// runtime index would copy it to scratch memory).
__device__ __forceinline__ int net_dim(const MlpNet &net, int l) {
  int v = net.dims[0];
#pragma unroll
  for (int j = 1; j <= MLP_MAXL; ++j)
    if (j == l) v = net.dims[j];
  return v;
}
__device__ __forceinline__ const __bf16 *net_weights(const MlpNet &net, int l) {
  const __bf16 *v = net.Wp[0];
#pragma unroll
  for (int j = 1; j < MLP_MAXL; ++j)
    if (j == l) v = net.Wp[j];
  return v;
}
__device__ __forceinline__ const float *net_bias(const MlpNet &net, int l) {
  const float *v = net.bias[0];
#pragma unroll
  for (int j = 1; j < MLP_MAXL; ++j)
    if (j == l) v = net.bias[j];
  return v;
}

// NB: 16-column fragments per wave (N = 128 NB); KT: the input width when it
// is known at compile time (the k loop is then unrolled whole, which spares
// the loop header's wait for every outstanding load and the loads past the
// last slice), else 0.
template <int NB, int KT>
__device__ __forceinline__ void mlp_layer(const __bf16 *__restrict__ Wp, const float *__restrict__ bias,
                                          int Kr, bool last, char *Hs, const float (*lnb)[4],
                                          int m0, int M, float *__restrict__ out, int ldo) {
  constexpr int N = 128 * NB;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int K = KT ? KT : Kr, n0 = wave * 16 * NB;
  // fragment c of the slice at k: Wp[k N + lane_off + 512 c .. +8]
  const uint32_t lane_off = (uint32_t)(n0 + (lane & 15)) * 32u + (uint32_t)(lane >> 4) * 8u;
  auto wload = [&](bf16x8 &dst, int c, int k) {
    // a wave-uniform base (scalar registers) plus the lane's fixed offset,
    // so the unrolled loop needs no vector register per address
    const __bf16 *base = Wp + ((size_t)k * N + (size_t)c * 512);
    dst = *(const bf16x8 *)(base + lane_off);
  };
  // Row fragment r at k: byte offset hb[r / 4] + (16 (r % 4) + k / 32) 1024
  // in Hs.  Rows 64.. get a base register of their own (kept opaque to the
  // compiler) so that every offset fits a DS instruction's 16-bit immediate:
  // one address register for the unrolled loop instead of one per load.
  const uint32_t hb0 = (uint32_t)lane * 16u;
  uint32_t hb1 = hb0 + 4u * 16u * 1024u;
  asm volatile("" : "+v"(hb1));
  auto hload = [&](bf16x8 &dst, int r, int k) {
    const uint32_t off = (r < 4 ? hb0 : hb1) + (uint32_t)((r & 3) * 16 + (k >> 5)) * 1024u;
    dst = *(const bf16x8 *)(Hs + off);
  };
  floatx4 acc[NB][8];
#pragma unroll
  for (int c = 0; c < NB; ++c)
#pragma unroll
    for (int r = 0; r < 8; ++r) acc[c][r] = floatx4{0.f, 0.f, 0.f, 0.f};
  // Software pipeline, fixed with scheduling barriers (left alone, the
  // compiler sinks each load to just before its use): row fragment r of the
  // next slice loads as soon as the MFMAs of this slice have read slot r, and
  // weight fragments load after the first row block of a slice into a set of
  // registers of their own: for the slice after next when the k loop is
  // unrolled (three sets), for the next slice otherwise (two sets).
  bf16x8 h[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) hload(h[r], r, 0);
  if constexpr (KT != 0) {
    constexpr int NS = KT / 32;
    bf16x8 w[3][NB];
#pragma unroll
    for (int c = 0; c < NB; ++c) {
      wload(w[0][c], c, 0);
      if (NS > 1) wload(w[1][c], c, 32);
    }
#pragma unroll
    for (int sl = 0; sl < NS; ++sl) {
      // no loads past the last slice (compile-time conditions: nothing is
      // left in flight for the barrier at the end of the layer to wait for)
#pragma unroll
      for (int r = 0; r < 8; ++r) {
#pragma unroll
        for (int c = 0; c < NB; ++c)
          acc[c][r] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w[sl % 3][c], h[r], acc[c][r], 0, 0, 0);
        if (sl + 1 < NS) hload(h[r], r, (sl + 1) * 32);
        if (r == 0 && sl + 2 < NS) {
#pragma unroll
          for (int c = 0; c < NB; ++c) wload(w[(sl + 2) % 3][c], c, (sl + 2) * 32);
        }
        __builtin_amdgcn_sched_barrier(0);
      }
    }
  } else {
    bf16x8 wa[NB], wb[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c) wload(wa[c], c, 0);
    auto slice = [&](const bf16x8(&w)[NB], bf16x8(&wn)[NB], int kn) {
#pragma unroll
      for (int r = 0; r < 8; ++r) {
#pragma unroll
        for (int c = 0; c < NB; ++c) acc[c][r] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w[c], h[r], acc[c][r], 0, 0, 0);
        hload(h[r], r, kn);
        if (r == 0) {
#pragma unroll
          for (int c = 0; c < NB; ++c) wload(wn[c], c, kn);
        }
        __builtin_amdgcn_sched_barrier(0);
      }
    };
    for (int k0 = 0; k0 < K; k0 += 64) { // K is a multiple of 64
      slice(wa, wb, k0 + 32);
      // the last slice's loads fetch that slice again (no branch in the loop)
      slice(wb, wa, k0 + 64 < K ? k0 + 64 : k0 + 32);
    }
  }
  __syncthreads(); // every wave has read this layer's input
  // The epilogue's indices derive from an opaque copy of the lane id: left
  // alone, the compiler computes them for every layer variant up front and
  // keeps them in scratch memory across the k loops (spill traffic in HBM).
  int le = lane;
  asm volatile("" : "+v"(le));
#pragma unroll
  for (int c = 0; c < NB; ++c) {
    const int n = n0 + c * 16 + (le >> 4) * 4;
    const floatx4 b = *(const floatx4 *)&bias[n];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int m = r * 16 + (le & 15);
      floatx4 v;
#pragma unroll
      for (int q = 0; q < 4; ++q) v[q] = acc[c][r][q] + b[q];
      if (!last) {
        __bf16 hv[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) hv[q] = (__bf16)(v[q] > 0.f ? v[q] : 0.f);
        __builtin_memcpy(Hs + hs_offset(m, n), hv, sizeof hv);
      } else {
        if (lnb) {
          // head_offsets_kernel: + ln bound[s] on the negative-binomial log
          // means, columns s*3K + K + c for c < K/2
          constexpr int KC = REBOP_SG_COMPONENTS;
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int s = (n + q) / (3 * KC), j = (n + q) % (3 * KC);
            if (s < 4 && j >= KC && j < KC + KC / 2) v[q] += lnb[m][s];
          }
        }
        if constexpr (NB == 1) {
          // a head of 128 columns goes through LDS (free now) so that the
          // stores below write whole rows
          *(floatx4 *)(Hs + m * MLP_HEAD_LD + n * 4) = v;
        } else if (m0 + m < M) {
          __builtin_nontemporal_store(v, (floatx4 *)&out[(size_t)(m0 + m) * ldo + n]);
        }
      }
    }
  }
  __syncthreads(); // the layer's output is whole before the next layer reads it
  if (NB == 1 && last) {
    // 128 rows x 512 B: a wave instruction stores two whole rows (full
    // cache lines, streaming: the head is not read again, and
    // write-allocated lines would push the weights out of L2)
    int te = tid;
    asm volatile("" : "+v"(te));
#pragma unroll
    for (int i = 0; i < MLP_BM * 32 / 512; ++i) {
      const int idx = te + 512 * i, m = idx >> 5, c4 = idx & 31;
      if (m0 + m < M)
        __builtin_nontemporal_store(*(const floatx4 *)(Hs + m * MLP_HEAD_LD + c4 * 16),
                                    (floatx4 *)&out[(size_t)(m0 + m) * ldo + c4 * 4]);
    }
  }
}

template <bool FEAT>
__global__ void __launch_bounds__(512) mlp_infer_kernel(int M, const __bf16 *__restrict__ X, int ldx,
                                                        MlpItems items, MlpNet net,
                                                        float *__restrict__ out, int ldo) {
  __shared__ __attribute__((aligned(16))) char Hs[MLP_BM * MLP_MAXW * 2];
  __shared__ float lnb[FEAT ? MLP_BM : 1][4];
  __shared__ double lg[FEAT ? MLP_BM : 1][8]; // the items' logs, shared by their four threads
  __shared__ __attribute__((aligned(16))) float Bs[MLP_MAXL * MLP_MAXW]; // every layer's bias
  const int tid = threadIdx.x, m0 = blockIdx.x * MLP_BM, K0 = net.dims[0];
  // every layer's bias: loaded into registers here, stored to LDS after the
  // inputs, so its latency overlaps theirs
  float bv[MLP_MAXL];
#pragma unroll
  for (int l = 0; l < MLP_MAXL; ++l)
    bv[l] = l < net.depth && tid < net.dims[l + 1] ? net.bias[l][tid] : 0.f;
  if (FEAT) {
    // four threads per item: thread q computes logs 2q, 2q+1 (of seven),
    // which the item's threads share through LDS, then features 4q .. 4q+3
    // (the arithmetic of item_features, so the bits of features_kernel)
    static_assert(REBOP_SG_FEATURES == 16 && MLP_BM * 4 == 512, "four threads x four features per item");
    const int item = tid >> 2, q = tid & 3;
    ItemInputs v{1.0, 1.0, 1.0, 1.0, 1.0, 1.0};
    if (m0 + item < M)
      v = item_inputs(items.item_point[m0 + item], items.item_sample[m0 + item], items.rates, items.x0,
                      items.tmax, items.nb_steps);
    lg[item][2 * q] = item_log(v, 2 * q);
    if (q < 3) lg[item][2 * q + 1] = item_log(v, 2 * q + 1);
    __syncthreads();
    double l7[7];
#pragma unroll
    for (int j = 0; j < 7; ++j) l7[j] = lg[item][j];
    __bf16 h4[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const float f = m0 + item < M ? (float)item_feature(v, l7, items.tmax, 4 * q + j) : 0.f;
      h4[j] = (__bf16)f;
    }
    __builtin_memcpy(Hs + hs_offset(item, 4 * q), h4, sizeof h4);
    const __bf16 z4[4] = {};
    for (int c = REBOP_SG_FEATURES + 4 * q; c < K0; c += 16) __builtin_memcpy(Hs + hs_offset(item, c), z4, sizeof z4);
    if (q == 0) {
      const float le = logf((float)v.e0), ls = logf((float)v.s0);
      lnb[item][0] = le;
      lnb[item][1] = ls;
      lnb[item][2] = le;
      lnb[item][3] = ls;
    }
  } else {
    // rows m0.. of X, 8 columns per thread and pass; rows past M read as zero
    const int per_row = K0 / 8;
    for (int idx = tid; idx < MLP_BM * per_row; idx += 512) {
      const int r = idx / per_row, c = (idx % per_row) * 8;
      bf16x8 v = {};
      if (m0 + r < M) v = *(const bf16x8 *)&X[(size_t)(m0 + r) * ldx + c];
      *(bf16x8 *)(Hs + hs_offset(r, c)) = v;
    }
  }
#pragma unroll
  for (int l = 0, o = 0; l < MLP_MAXL; o += net.dims[l + 1], ++l)
    if (l < net.depth && tid < net.dims[l + 1]) Bs[o + tid] = bv[l];
  __syncthreads();
  const float(*lp)[4] = FEAT ? lnb : nullptr;
  for (int l = 0, o = 0; l < net.depth; o += net_dim(net, l + 1), ++l) {
    const __bf16 *W = net_weights(net, l);
    const float *bias = Bs + o;
    const int K = net_dim(net, l), N = net_dim(net, l + 1);
    const bool last = l == net.depth - 1;
    // the surrogate's shapes, fully unrolled
    if (K == 512 && N == 512) {
      mlp_layer<4, 512>(W, bias, K, last, Hs, lp, m0, M, out, ldo);
      continue;
    }
    if (K == 64 && N == 512) {
      mlp_layer<4, 64>(W, bias, K, last, Hs, lp, m0, M, out, ldo);
      continue;
    }
    if (K == 512 && N == 128) {
      mlp_layer<1, 512>(W, bias, K, last, Hs, lp, m0, M, out, ldo);
      continue;
    }
    switch (N) {
    case 128: mlp_layer<1, 0>(W, bias, K, last, Hs, lp, m0, M, out, ldo); break;
    case 256: mlp_layer<2, 0>(W, bias, K, last, Hs, lp, m0, M, out, ldo); break;
    default: mlp_layer<4, 0>(W, bias, K, last, Hs, lp, m0, M, out, ldo); break;
    }
  }
}

// ---- special functions (FP64): recurrence to z >= 10, then the asymptotic
// series of ln Gamma (Stirling) and of the digamma function, with the
// Bernoulli-number coefficients B_2n / (2n (2n-1)) and B_2n / (2n).
__device__ __forceinline__ double lgamma_pos(double z) {
  // ln Gamma(z) = ln Gamma(z + n) - ln(z (z+1) ... (z+n-1)); one log of the
  // product (at most 10 factors below 10: no overflow)
  double prod = 1.0;
  while (z < 10.0) {
    prod *= z;
    z += 1.0;
  }
  const double shift = log(prod);
  const double iz = 1.0 / z, iz2 = iz * iz;
  const double series =
      iz * (1.0 / 12.0 - iz2 * (1.0 / 360.0 - iz2 * (1.0 / 1260.0 - iz2 * (1.0 / 1680.0 - iz2 * (1.0 / 1188.0)))));
  return (z - 0.5) * log(z) - z + 0.91893853320467274178 + series - shift;
}

__device__ __forceinline__ double digamma_pos(double z) {
  double shift = 0.0;
  while (z < 10.0) {
    shift += 1.0 / z;
    z += 1.0;
  }
  const double iz = 1.0 / z, iz2 = iz * iz;
  const double series =
      iz2 * (1.0 / 12.0 - iz2 * (1.0 / 120.0 - iz2 * (1.0 / 252.0 - iz2 * (1.0 / 240.0 - iz2 * (1.0 / 132.0)))));
  return log(z) - 0.5 * iz - series - shift;
}

// Negative log-likelihood of the SSA samples under the predicted mixtures,
// and its gradient w.r.t. the head outputs.  One wave per item, one sample
// per lane (64 samples).  Head layout per species s (K components):
// [s*3K + c] mixture logit, [s*3K + K + c] log mean, [s*3K + 2K + c] log
// dispersion.  With bounds, components c < K/2 are negative binomials of the
// count x with log means relative to ln bound, and components c >= K/2 are
// binomial(bound, p) with logit p in the log-mean slot (a bounded count can
// vary less than its mean, which no negative binomial does; their dispersion
// slot is unused).  samples: the SSA output
// [nb+1][S][P * J], item (p, i) reads samples[(i*S + s) * P*J + p*J + j].
// grad = d(mean NLL over items and samples)/d head (FP32, [B][ldh]);
// loss_sum += total NLL (FP64).
template <int K>
__global__ void __launch_bounds__(64) nb_mixture_nll_kernel(
    int B, int S, int J, const float *__restrict__ head, int ldh,
    const uint32_t *__restrict__ item_point, const uint32_t *__restrict__ item_sample,
    const int64_t *__restrict__ samples, uint64_t n_traj_total, float *__restrict__ grad,
    double *__restrict__ loss_sum, double scale, const float *__restrict__ bound) {
  const int b = blockIdx.x;
  const int j = threadIdx.x;
  if (b >= B) return;
  const uint32_t p = item_point[b], i = item_sample[b];
  const bool active = j < J;
  double lane_loss = 0.0;
  for (int s = 0; s < S; ++s) {
    const float *h = head + (size_t)b * ldh + (size_t)s * 3 * K;
    const double xs = active ? (double)samples[((size_t)i * S + s) * n_traj_total + (size_t)p * J + j] : 0.0;
    const double bnd = bound ? (double)bound[(size_t)b * S + s] : 0.0;
    const double off = bound ? log(bnd) : 0.0;
    // log softmax of the logits (same on every lane)
    double lmax = -1e300;
#pragma unroll
    for (int c = 0; c < K; ++c) lmax = fmax(lmax, (double)h[c]);
    double zs = 0.0;
#pragma unroll
    for (int c = 0; c < K; ++c) zs += exp((double)h[c] - lmax);
    const double lz = lmax + log(zs);
    double lterm[K], mu[K], r[K], lr[K], lmu[K];
    const double lgx1 = lgamma_pos(xs + 1.0);
    // ln C(bound, x) for the binomial components of a bounded species
    const double lcb = bound ? lgamma_pos(bnd + 1.0) - lgx1 - lgamma_pos(fmax(bnd - xs, 0.0) + 1.0) : 0.0;
    double best = -1e300;
#pragma unroll
    for (int c = 0; c < K; ++c) {
      double ll;
      if (bound && c >= K / 2) {
        // binomial(bound, p), p = sigmoid(theta): ln p = -softplus(-theta)
        lmu[c] = fmin(fmax((double)h[K + c], -30.0), 30.0);
        const double th = lmu[c];
        const double sp_neg = th < 0.0 ? log1p(exp(th)) - th : log1p(exp(-th)); // softplus(-th)
        const double sp_pos = sp_neg + th;                                          // softplus(th)
        mu[c] = exp(-sp_neg); // p
        ll = lcb - xs * sp_neg - (bnd - xs) * sp_pos;
      } else {
        lmu[c] = fmin(fmax((double)h[K + c] + off, -8.0), 9.0);
        lr[c] = fmin(fmax((double)h[2 * K + c], -6.0), 12.0);
        mu[c] = exp(lmu[c]);
        r[c] = exp(lr[c]);
        const double lrm = log(r[c] + mu[c]);
        ll = lgamma_pos(xs + r[c]) - lgamma_pos(r[c]) - lgx1 + r[c] * (lr[c] - lrm) +
             xs * (lmu[c] - lrm);
      }
      lterm[c] = ((double)h[c] - lz) + ll;
      best = fmax(best, lterm[c]);
    }
    double se = 0.0;
#pragma unroll
    for (int c = 0; c < K; ++c) se += exp(lterm[c] - best);
    const double lp = best + log(se);
    if (active) lane_loss -= lp;
#pragma unroll
    for (int c = 0; c < K; ++c) {
      const double gam = active ? exp(lterm[c] - lp) : 0.0;
      const double pi = exp((double)h[c] - lz);
      double gl = active ? (pi - gam) : 0.0;
      double gm, gr;
      if (bound && c >= K / 2) {
        gm = -gam * (xs - bnd * mu[c]); // d/dtheta of -ln binomial
        gr = 0.0;
        if ((double)h[K + c] != lmu[c]) gm = 0.0; // clamped
      } else {
        const double rm = r[c] + mu[c];
        gm = -gam * r[c] * (xs - mu[c]) / rm;
        gr = -gam * r[c] *
             (digamma_pos(xs + r[c]) - digamma_pos(r[c]) + (lr[c] - log(rm)) + (mu[c] - xs) / rm);
        if ((double)h[K + c] + off != lmu[c]) gm = 0.0; // clamped
        if ((double)h[2 * K + c] != lr[c]) gr = 0.0;
      }
      // sum over the wave's samples
#pragma unroll
      for (int o = 32; o >= 1; o >>= 1) {
        gl += __shfl_xor(gl, o);
        gm += __shfl_xor(gm, o);
        gr += __shfl_xor(gr, o);
      }
      if (j == 0) {
        float *g = grad + (size_t)b * ldh + (size_t)s * 3 * K;
        g[c] = (float)(gl * scale);
        g[K + c] = (float)(gm * scale);
        g[2 * K + c] = (float)(gr * scale);
      }
    }
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) lane_loss += __shfl_xor(lane_loss, off);
  if (j == 0) atomicAdd(loss_sum, lane_loss);
}

} // namespace sg

// ---------------------------------------------------------------------------
// C ABI

namespace {
hipError_t g_err = hipSuccess;
int ret(hipError_t e) {
  g_err = e;
  return e == hipSuccess ? 0 : 4;
}
unsigned nblk(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

// The layer table of the fused forward pass; 3 if the shapes are outside
// what mlp_infer_kernel handles.
int mlp_net(int depth, const int *dims, const void *const *Wp, const float *const *bias,
            sg::MlpNet &net) {
  if (depth < 1 || depth > sg::MLP_MAXL || !dims || !Wp || !bias) return 3;
  if (dims[0] <= 0 || dims[0] % 64 || dims[0] > sg::MLP_MAXW) return 3;
  net.depth = depth;
  net.dims[0] = dims[0];
  for (int l = 0; l < depth; ++l) {
    const int n = dims[l + 1];
    if (n != 128 && n != 256 && n != 512) return 3;
    if (!Wp[l] || !bias[l] || (uintptr_t)Wp[l] % 16 || (uintptr_t)bias[l] % 16) return 3;
    net.dims[l + 1] = n;
    net.Wp[l] = (const __bf16 *)Wp[l];
    net.bias[l] = bias[l];
  }
  return 0;
}

template <bool TA, bool TB>
void launch_gemm(bool big, dim3 grid, hipStream_t s, int M, int N, int K, const __bf16 *a,
                        int lda, const __bf16 *b, int ldb, float *C, int ldc,
                        const sg::Epilogue *epi = nullptr) {
  const sg::Epilogue none{nullptr, 0, nullptr, 0, nullptr, 0};
  if (!TA && TB) { // k-contiguous operands: the pipelined kernel
    if (big && epi)
      hipLaunchKernelGGL((sg::gemm_nt_kernel<128, 128, true>), grid, dim3(256), 0, s, M, N, K, a, lda,
                         b, ldb, C, ldc, *epi);
    else if (epi)
      hipLaunchKernelGGL((sg::gemm_nt_kernel<64, 64, true>), grid, dim3(256), 0, s, M, N, K, a, lda,
                         b, ldb, C, ldc, *epi);
    else if (big)
      hipLaunchKernelGGL((sg::gemm_nt_kernel<128, 128, false>), grid, dim3(256), 0, s, M, N, K, a,
                         lda, b, ldb, C, ldc, none);
    else
      hipLaunchKernelGGL((sg::gemm_nt_kernel<64, 64, false>), grid, dim3(256), 0, s, M, N, K, a, lda,
                         b, ldb, C, ldc, none);
    return;
  }
  if (epi && big)
    hipLaunchKernelGGL((sg::gemm_kernel<TA, TB, 128, 128, true>), grid, dim3(256), 0, s, M, N, K, a,
                       lda, b, ldb, C, ldc, *epi);
  else if (epi)
    hipLaunchKernelGGL((sg::gemm_kernel<TA, TB, 64, 64, true>), grid, dim3(256), 0, s, M, N, K, a,
                       lda, b, ldb, C, ldc, *epi);
  else if (big)
    hipLaunchKernelGGL((sg::gemm_kernel<TA, TB, 128, 128, false>), grid, dim3(256), 0, s, M, N, K, a,
                       lda, b, ldb, C, ldc, none);
  else
    hipLaunchKernelGGL((sg::gemm_kernel<TA, TB, 64, 64, false>), grid, dim3(256), 0, s, M, N, K, a,
                       lda, b, ldb, C, ldc, none);
}

} // namespace

extern "C" {

const char *rebop_sg_last_error(void) { return hipGetErrorString(g_err); }

int rebop_sg_gemm_bf16(int ta, int tb, int M, int N, int K, const void *A, int lda, const void *B,
                       int ldb, float *C, int ldc, void *stream) {
  if (M % 64 || N % 64 || K % 32 || M <= 0 || N <= 0 || K <= 0) return 3;
  // 128 x 128 tiles when there are enough of them to fill the chip
  const bool big = M % 128 == 0 && N % 128 == 0 && (M / 128) * (N / 128) >= 512;
  const int tile = big ? 128 : 64;
  // split K when the output has too few tiles to fill the chip
  int splits = 1;
  while (!big && (M / 64) * (N / 64) * splits < 512 && K % (64 * splits) == 0 &&
         K / (32 * splits) >= 8)
    splits *= 2;
  const dim3 grid(N / tile, M / tile, splits);
  const __bf16 *a = (const __bf16 *)A, *b = (const __bf16 *)B;
  hipStream_t s = (hipStream_t)stream;
  if (splits > 1)
    hipLaunchKernelGGL(sg::zero_kernel, dim3(nblk((size_t)M * N, 256)), dim3(256), 0, s, M, N, C, ldc);
  if (!ta && !tb) launch_gemm<false, false>(big, grid, s, M, N, K, a, lda, b, ldb, C, ldc);
  else if (!ta && tb) launch_gemm<false, true>(big, grid, s, M, N, K, a, lda, b, ldb, C, ldc);
  else if (ta && !tb) launch_gemm<true, false>(big, grid, s, M, N, K, a, lda, b, ldb, C, ldc);
  else launch_gemm<true, true>(big, grid, s, M, N, K, a, lda, b, ldb, C, ldc);
  return ret(hipGetLastError());
}

int rebop_sg_gemm_bias_act_bf16(int tb, int M, int N, int K, const void *A, int lda, const void *B,
                                int ldb, const float *bias, int relu, void *H, int ldh, float *F,
                                int ldf, void *stream) {
  if (M % 64 || N % 64 || K % 32 || M <= 0 || N <= 0 || K <= 0) return 3;
  if (!bias || (!H && !F)) return 3;
  const bool big = M % 128 == 0 && N % 128 == 0 && (M / 128) * (N / 128) >= 512;
  const int tile = big ? 128 : 64;
  const dim3 grid(N / tile, M / tile, 1); // no split K: the epilogue needs the whole sum
  const sg::Epilogue epi{bias, relu, (__bf16 *)H, ldh, F, ldf};
  if (tb)
    launch_gemm<false, true>(big, grid, (hipStream_t)stream, M, N, K, (const __bf16 *)A, lda,
                             (const __bf16 *)B, ldb, nullptr, 0, &epi);
  else
    launch_gemm<false, false>(big, grid, (hipStream_t)stream, M, N, K, (const __bf16 *)A, lda,
                              (const __bf16 *)B, ldb, nullptr, 0, &epi);
  return ret(hipGetLastError());
}

int rebop_sg_bias_act(int M, int N, const float *C, int ldc, const float *bias, int relu, void *H,
                      int ldh, float *F, int ldf, void *stream) {
  hipLaunchKernelGGL(sg::bias_act_kernel, dim3(nblk((size_t)M * N, 256)), dim3(256), 0,
                     (hipStream_t)stream, M, N, C, ldc, bias, relu, (__bf16 *)H, ldh, F, ldf);
  return ret(hipGetLastError());
}

int rebop_sg_relu_grad(int M, int N, float *dH, int ldd, const void *H, int ldh, void *dZ, int ldz,
                       void *stream) {
  hipLaunchKernelGGL(sg::relu_grad_kernel, dim3(nblk((size_t)M * N, 256)), dim3(256), 0,
                     (hipStream_t)stream, M, N, dH, ldd, (const __bf16 *)H, ldh, (__bf16 *)dZ, ldz);
  return ret(hipGetLastError());
}

int rebop_sg_transpose_bf16(int rows, int cols, const float *x, void *y, void *stream) {
  if (rows <= 0 || cols <= 0) return 3;
  hipLaunchKernelGGL(sg::transpose_bf16_kernel, dim3(nblk((size_t)rows * cols, 256)), dim3(256), 0,
                     (hipStream_t)stream, rows, cols, x, (__bf16 *)y);
  return ret(hipGetLastError());
}

int rebop_sg_cast_bf16(size_t n, const float *x, void *y, void *stream) {
  hipLaunchKernelGGL(sg::cast_bf16_kernel, dim3(nblk(n, 256)), dim3(256), 0, (hipStream_t)stream,
                     n, x, (__bf16 *)y);
  return ret(hipGetLastError());
}

int rebop_sg_colsum(int M, int N, const float *X, int ldx, float *out, void *stream) {
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(sg::fill_kernel, dim3(nblk((size_t)N, 256)), dim3(256), 0, st, N, out);
  hipLaunchKernelGGL(sg::colsum_kernel, dim3(nblk((size_t)N, 64), nblk((size_t)M, 256)), dim3(256),
                     0, st, M, N, X, ldx, out);
  return ret(hipGetLastError());
}

int rebop_sg_adam(size_t n, float *w, const float *g, float *m, float *v, void *wb, float lr,
                  float b1, float b2, float eps, float bc1, float bc2, void *stream) {
  hipLaunchKernelGGL(sg::adam_kernel, dim3(nblk(n, 256)), dim3(256), 0, (hipStream_t)stream, n, w,
                     g, m, v, (__bf16 *)wb, lr, b1, b2, eps, bc1, bc2);
  return ret(hipGetLastError());
}

int rebop_sg_features(int B, const uint32_t *item_point, const uint32_t *item_sample,
                      const double *rates, const int64_t *x0, double tmax, int nb_steps, void *X,
                      int ldx, float *bound, void *stream) {
  if (ldx < REBOP_SG_FEATURES) return 3;
  hipLaunchKernelGGL(sg::features_kernel, dim3(nblk((size_t)B, 256)), dim3(256), 0,
                     (hipStream_t)stream, B, item_point, item_sample, rates, x0, tmax, nb_steps,
                     (__bf16 *)X, ldx, bound);
  return ret(hipGetLastError());
}

int rebop_sg_head_offsets(int B, int S, int K, float *head, int ldh, const float *bound,
                          void *stream) {
  if (ldh < S * 3 * K) return 3;
  hipLaunchKernelGGL(sg::head_offsets_kernel, dim3(nblk((size_t)B * S * K, 256)), dim3(256), 0,
                     (hipStream_t)stream, B, S, K, head, ldh, bound);
  return ret(hipGetLastError());
}

int rebop_sg_pack_kslices_bf16(int K, int N, const float *W, void *Wp, void *stream) {
  if (K <= 0 || N <= 0 || K % 32 || !W || !Wp) return 3;
  hipLaunchKernelGGL(sg::pack_kslices_kernel, dim3(nblk((size_t)K * N, 256)), dim3(256), 0,
                     (hipStream_t)stream, K, N, W, (__bf16 *)Wp);
  return ret(hipGetLastError());
}

int rebop_sg_mlp_forward_bf16(int M, const void *X, int ldx, int depth, const int *dims,
                              const void *const *Wp, const float *const *bias, float *out, int ldo,
                              void *stream) {
  sg::MlpNet net{};
  if (mlp_net(depth, dims, Wp, bias, net) || M <= 0 || !X || !out || (uintptr_t)X % 16 ||
      (uintptr_t)out % 16 || ldx < dims[0] || ldx % 8 || ldo < dims[depth] || ldo % 4)
    return 3;
  const sg::MlpItems none{nullptr, nullptr, nullptr, nullptr, 0.0, 1};
  hipLaunchKernelGGL(sg::mlp_infer_kernel<false>, dim3(nblk((size_t)M, sg::MLP_BM)), dim3(512), 0,
                     (hipStream_t)stream, M, (const __bf16 *)X, ldx, none, net, out, ldo);
  return ret(hipGetLastError());
}

int rebop_sg_infer(int B, const uint32_t *item_point, const uint32_t *item_sample,
                   const double *rates, const int64_t *x0, double tmax, int nb_steps, int depth,
                   const int *dims, const void *const *Wp, const float *const *bias, float *head,
                   int ldh, void *stream) {
  sg::MlpNet net{};
  if (mlp_net(depth, dims, Wp, bias, net) || B <= 0 || !item_point || !item_sample || !rates ||
      !x0 || !head || (uintptr_t)head % 16 || dims[0] < REBOP_SG_FEATURES ||
      dims[depth] < 4 * 3 * REBOP_SG_COMPONENTS || ldh < dims[depth] || ldh % 4 || nb_steps <= 0)
    return 3;
  const sg::MlpItems items{item_point, item_sample, rates, x0, tmax, nb_steps};
  hipLaunchKernelGGL(sg::mlp_infer_kernel<true>, dim3(nblk((size_t)B, sg::MLP_BM)), dim3(512), 0,
                     (hipStream_t)stream, B, nullptr, 0, items, net, head, ldh);
  return ret(hipGetLastError());
}

int rebop_sg_nb_mixture_nll(int B, int S, int K, int J, const float *head, int ldh,
                            const uint32_t *item_point, const uint32_t *item_sample,
                            const int64_t *samples, uint64_t n_traj_total, float *grad,
                            double *loss_sum, double scale, const float *bound, void *stream) {
  if (K != REBOP_SG_COMPONENTS || J > 64 || J <= 0) return 3;
  hipLaunchKernelGGL((sg::nb_mixture_nll_kernel<REBOP_SG_COMPONENTS>), dim3(B), dim3(64), 0,
                     (hipStream_t)stream, B, S, J, head, ldh, item_point, item_sample, samples,
                     n_traj_total, grad, loss_sum, scale, bound);
  return ret(hipGetLastError());
}

} // extern "C"
