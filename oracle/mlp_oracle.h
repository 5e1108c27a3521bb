/*
 * CPU oracle of the learned proposal network (FP64).
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the HIP MLP kernels; only tests,
 * the smoke check and the bench's CPU leg load it.  The reference has no
 * learned proposal; this file has no reference counterpart.
 *
 * Network: 22 -> 64 -> 64 -> 4, ReLU, parameters packed as
 *   W1[64][22] b1[64] W2[64][64] b2[64] W3[4][64] b3[4]   (5892 values).
 * Features of a site (orc_mlp_features): the six neighbour directors in the
 * order +x,-x,+y,-y,+z,-z after sign canonicalisation (the +x neighbour is
 * flipped to the upper half space, the others to a non-negative dot with it),
 * the field vector elecE*e, and 1/T.  Head: mu = normalize(base + o[0..2])
 * with base the sum of the canonical neighbours, log kappa = clamp(o[3],-6,8).
 * Proposal density (axial von Mises-Fisher mixture):
 *   q(n) = kappa / (4 pi sinh kappa) * cosh(kappa mu.n).
 */
#ifndef MLP_ORACLE_H_
#define MLP_ORACLE_H_
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MLP_IN 22
#define ORC_MLP_H 64
#define ORC_MLP_OUT 4
#define ORC_MLP_NPARAM (64 * 22 + 64 + 64 * 64 + 64 + 4 * 64 + 4)
#define ORC_LOGK_MIN (-6.0)
#define ORC_LOGK_MAX (8.0)

/* FP32 features and base axis, bit-identical to the kernels' own. */
void orc_mlp_features(const float nbr[18], const float field[3], float T, float x[22], float base[3]);

/* Round-to-nearest-even to bfloat16 (returned as float). */
float orc_bf16(float v);

/* Forward pass in FP64.  bf16 != 0 rounds the input and both hidden layers
 * to bfloat16 before each product, as the MFMA kernels feed them. */
void orc_mlp_forward(const float *params, const float x[22], int bf16, double out[4]);
void orc_mlp_head(const double out[4], const float base[3], double mu[3], double *logk);

/* Mean negative log-likelihood of the directors n (B x 3) under the proposal
 * of their neighbourhoods (X: B x 22, base: B x 3) and its gradient with
 * respect to the packed parameters (grad: 5892 values, overwritten). */
double orc_mlp_nll_grad(const float *params, int B, const float *X, const float *base, const float *n, int bf16,
                        double *grad);

/* log q(n) of the axial vMF mixture (FP64, stable for large kappa). */
double orc_vmf_logq(const double mu[3], double kappa, const double n[3]);

#ifdef __cplusplus
}
#endif
#endif
