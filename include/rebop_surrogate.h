/*
 * rebop_surrogate.h -- C ABI of the surrogate's device kernels.
 *
 * No counterpart in the reference (rebop has no learned component).  These
 * are the building blocks the training and inference loop (Python,
 * rebop_amd/surrogate.py) calls: a BF16 matrix-core GEMM with FP32
 * accumulation, bias/activation epilogues, the ReLU derivative, column sums,
 * Adam, the input featurisation of (parameter point, sample time) items, and
 * the negative-binomial-mixture likelihood head with its gradient.  Every
 * pointer is a device pointer; `stream` is a hipStream_t (NULL = default).
 * Functions return 0, 3 (invalid shape) or 4 (HIP error; see
 * rebop_sg_last_error).
 */
#ifndef REBOP_SURROGATE_H
#define REBOP_SURROGATE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define REBOP_SG_COMPONENTS 8 /* mixture components per species */
#define REBOP_SG_FEATURES 16   /* input features per item */

const char *rebop_sg_last_error(void);

/* C[M][N] (FP32, ldc) = op(A) op(B) with BF16 A and B.  ta: A is stored
 * [K][M] (else [M][K]); tb: B is stored [N][K] (else [K][N]).  M and N
 * multiples of 64, K a multiple of 32; leading dimensions in elements,
 * multiples of 8.  When the output has fewer than 512 tiles of 64 x 64, K is
 * split over several workgroups whose partial products are added into C
 * atomically (so the FP32 summation order varies from run to run). */
int rebop_sg_gemm_bf16(int ta, int tb, int M, int N, int K, const void *A, int lda, const void *B,
                       int ldb, float *C, int ldc, void *stream);

/* H = bf16(act(A B + bias)) and/or F = act(A B + bias) (FP32) in one launch:
 * the product of rebop_sg_gemm_bf16 (A [M][K]; B [K][N], or with tb the
 * [N][K] layout, which loads faster)
 * with rebop_sg_bias_act's epilogue applied to its FP32 accumulators, without
 * the FP32 product going through memory.  Never splits K, so where
 * rebop_sg_gemm_bf16 does not split it either (outputs of at least 512
 * tiles) the values are the same bits.  M, N multiples of 64, K of 32;
 * bias [N]; H or F may be NULL (not both).  Returns 3 on a bad shape or
 * argument. */
int rebop_sg_gemm_bias_act_bf16(int tb, int M, int N, int K, const void *A, int lda, const void *B,
                                int ldb, const float *bias, int relu, void *H, int ldh, float *F,
                                int ldf, void *stream);

/* H = bf16(act(C + bias)) and/or F = act(C + bias) (either may be NULL). */
int rebop_sg_bias_act(int M, int N, const float *C, int ldc, const float *bias, int relu, void *H,
                      int ldh, float *F, int ldf, void *stream);

/* dH *= (H > 0) in place (FP32), and dZ = bf16(dH). */
int rebop_sg_relu_grad(int M, int N, float *dH, int ldd, const void *H, int ldh, void *dZ, int ldz,
                       void *stream);

int rebop_sg_cast_bf16(size_t n, const float *x, void *y, void *stream);

/* y[c][r] = bf16(x[r][c]): the transposed BF16 copy of an FP32 matrix
 * (rows x cols, row-major). */
int rebop_sg_transpose_bf16(int rows, int cols, const float *x, void *y, void *stream);

/* out[n] = sum over m of X[m][n] (partial sums over row blocks are added
 * atomically, so the FP32 summation order varies from run to run). */
int rebop_sg_colsum(int M, int N, const float *X, int ldx, float *out, void *stream);

/* One Adam step on n FP32 weights; wb (nullable) receives their BF16 copy.
 * bc1, bc2: the bias corrections 1 - b1^t and 1 - b2^t. */
int rebop_sg_adam(size_t n, float *w, const float *g, float *m, float *v, void *wb, float lr,
                  float b1, float b2, float eps, float bc1, float bc2, void *stream);

/* X[b][0..ldx) (BF16) = the REBOP_SG_FEATURES features of item b, scaled to
 * about [-1, 1], then zeros: for parameter point item_point[b] and the time of
 * sample item_sample[b] (t = tmax * i / nb_steps), log10 of the three rate
 * constants, E0, S0, t, their log10, and log10 of kcat t, kf E0 t, kf S0 t,
 * K_M / S0 (K_M = (kb + kcat) / kf), kb / kcat, E0 / S0 and kcat E0 t / S0.
 * bound (nullable, [B][4]) receives E0, S0, E0, S0: the counts that bound
 * species E, S, ES, P (E + ES = E0 and S + ES + P = S0).  ldx >=
 * REBOP_SG_FEATURES. */
int rebop_sg_features(int B, const uint32_t *item_point, const uint32_t *item_sample,
                      const double *rates, const int64_t *x0, double tmax, int nb_steps, void *X,
                      int ldx, float *bound, void *stream);

/* head[b][s*3K + K + c] += ln bound[b][s] for the negative-binomial
 * components c < K/2: turns the head into the mixtures' parameters (logits;
 * log means and log dispersions, or logit p for the binomial components). */
int rebop_sg_head_offsets(int B, int S, int K, float *head, int ldh, const float *bound,
                          void *stream);

/* Wp = the k-slice-major BF16 copy of W [K][N] (FP32, row-major) that the
 * fused forward pass reads: Wp[k / 32][n][k % 32] = bf16(W[k][n]), so the
 * weights of 16 columns over one 32-wide k slice are 1 KB of consecutive
 * bytes.  K a multiple of 32. */
int rebop_sg_pack_kslices_bf16(int K, int N, const float *W, void *Wp, void *stream);

/* The whole network's forward pass in one launch (inference): layer l maps
 * width dims[l] to dims[l+1] with weights Wp[l] in the layout of
 * rebop_sg_pack_kslices_bf16 (BF16) and bias[l] (FP32); layers 0..depth-2
 * apply ReLU and round to BF16, the last writes out[M][dims[depth]] (FP32,
 * ldo).  Same values, bit for bit, as rebop_sg_gemm_bias_act_bf16 applied
 * layer by layer, with the activations kept on chip.  depth 1..8; dims[0] a
 * multiple of 64 up to 512, the others 128, 256 or 512; X [M][ldx] BF16
 * (ldx >= dims[0], a multiple of 8); ldo >= dims[depth], a multiple of 4;
 * X, out, Wp[l] and bias[l] 16-byte aligned.  Wp and bias are host arrays of
 * device pointers, dims a host array. */
int rebop_sg_mlp_forward_bf16(int M, const void *X, int ldx, int depth, const int *dims,
                              const void *const *Wp, const float *const *bias, float *out, int ldo,
                              void *stream);

/* rebop_sg_features, rebop_sg_mlp_forward_bf16 and rebop_sg_head_offsets
 * (S = 4, K = REBOP_SG_COMPONENTS) in one launch: head[b] = the mixture
 * parameters of item b (point item_point[b], sample item_sample[b]), the same
 * bits as the three calls.  dims[0] >= REBOP_SG_FEATURES (inputs past the
 * features are zero), dims[depth] >= 12 K, ldh >= dims[depth]. */
int rebop_sg_infer(int B, const uint32_t *item_point, const uint32_t *item_sample,
                   const double *rates, const int64_t *x0, double tmax, int nb_steps, int depth,
                   const int *dims, const void *const *Wp, const float *const *bias, float *head,
                   int ldh, void *stream);

/* Negative log-likelihood of J SSA samples per item under the predicted
 * per-species K-component negative-binomial mixtures (K = REBOP_SG_COMPONENTS),
 * and its gradient: grad[b][...] = scale * d NLL / d head[b][...];
 * *loss_sum += NLL.  samples: an SSA output [nb+1][S][n_traj_total] in which
 * item (p, i) owns trajectories p*J .. p*J+J-1 of sample i.  With bound
 * ([B][S], nullable): components 0 .. K/2-1 are negative binomials whose log
 * means are relative to ln bound, and components K/2 .. K-1 are
 * binomial(bound, p) with logit p in the log-mean slot and an unused
 * dispersion slot (a bounded count can vary less than its mean, which no
 * negative binomial does). */
int rebop_sg_nb_mixture_nll(int B, int S, int K, int J, const float *head, int ldh,
                            const uint32_t *item_point, const uint32_t *item_sample,
                            const int64_t *samples, uint64_t n_traj_total, float *grad,
                            double *loss_sum, double scale, const float *bound, void *stream);

#ifdef __cplusplus
}
#endif

#endif
