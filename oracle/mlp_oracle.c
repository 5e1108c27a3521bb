/*
 * CPU oracle of the learned proposal network (see mlp_oracle.h).
 * TEST INFRASTRUCTURE ONLY.
 */
#include "mlp_oracle.h"

#include <math.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#define W1_OFF 0
#define B1_OFF (64 * 22)
#define W2_OFF (B1_OFF + 64)
#define B2_OFF (W2_OFF + 64 * 64)
#define W3_OFF (B2_OFF + 64)
#define B3_OFF (W3_OFF + 4 * 64)

void orc_mlp_features(const float nbr[18], const float field[3], float T, float x[22], float base[3]) {
  float c0[3] = {nbr[0], nbr[1], nbr[2]};
  int flip0 = c0[2] < 0.0f || (c0[2] == 0.0f && (c0[1] < 0.0f || (c0[1] == 0.0f && c0[0] < 0.0f)));
  if (flip0) {
    c0[0] = -c0[0];
    c0[1] = -c0[1];
    c0[2] = -c0[2];
  }
  x[0] = c0[0];
  x[1] = c0[1];
  x[2] = c0[2];
  base[0] = c0[0];
  base[1] = c0[1];
  base[2] = c0[2];
  for (int b = 1; b < 6; ++b) {
    const float *m = nbr + 3 * b;
    float d = c0[0] * m[0];
    d = d + c0[1] * m[1];
    d = d + c0[2] * m[2];
    const float s = d < 0.0f ? -1.0f : 1.0f;
    for (int c = 0; c < 3; ++c) {
      x[3 * b + c] = s * m[c];
      base[c] = base[c] + s * m[c];
    }
  }
  x[18] = field[0];
  x[19] = field[1];
  x[20] = field[2];
  x[21] = 1.0f / T;
}

float orc_bf16(float v) {
  uint32_t u;
  memcpy(&u, &v, 4);
  if ((u & 0x7F800000u) == 0x7F800000u) { /* inf / nan: keep, quiet */
    if (u & 0x007FFFFFu) u |= 0x00400000u;
    u &= 0xFFFF0000u;
  } else {
    const uint32_t lsb = (u >> 16) & 1u;
    u += 0x7FFFu + lsb;
    u &= 0xFFFF0000u;
  }
  float r;
  memcpy(&r, &u, 4);
  return r;
}

static inline double rnd(double v, int bf16) { return bf16 ? (double)orc_bf16((float)v) : v; }

/* Forward keeping the activations for the backward pass. */
static void forward_full(const float *P, const float x[22], int bf16, double xin[22], double pre1[64],
                         double h1in[64], double pre2[64], double h2in[64], double out[4]) {
  for (int k = 0; k < 22; ++k) xin[k] = rnd(x[k], bf16);
  for (int j = 0; j < 64; ++j) {
    double a = 0.0;
    for (int k = 0; k < 22; ++k) a += (double)P[W1_OFF + 22 * j + k] * xin[k];
    a += (double)P[B1_OFF + j];
    pre1[j] = a;
    h1in[j] = rnd(a > 0 ? a : 0.0, bf16);
  }
  for (int j = 0; j < 64; ++j) {
    double a = 0.0;
    for (int k = 0; k < 64; ++k) a += (double)P[W2_OFF + 64 * j + k] * h1in[k];
    a += (double)P[B2_OFF + j];
    pre2[j] = a;
    h2in[j] = rnd(a > 0 ? a : 0.0, bf16);
  }
  for (int o = 0; o < 4; ++o) {
    double a = 0.0;
    for (int k = 0; k < 64; ++k) a += (double)P[W3_OFF + 64 * o + k] * h2in[k];
    out[o] = a + (double)P[B3_OFF + o];
  }
}

void orc_mlp_forward(const float *params, const float x[22], int bf16, double out[4]) {
  double xin[22], p1[64], h1[64], p2[64], h2[64];
  forward_full(params, x, bf16, xin, p1, h1, p2, h2, out);
}

void orc_mlp_head(const double out[4], const float base[3], double mu[3], double *logk) {
  double v[3] = {base[0] + out[0], base[1] + out[1], base[2] + out[2]};
  double r = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  if (r == 0) r = 1;
  mu[0] = v[0] / r;
  mu[1] = v[1] / r;
  mu[2] = v[2] / r;
  double l = out[3];
  if (l < ORC_LOGK_MIN) l = ORC_LOGK_MIN;
  if (l > ORC_LOGK_MAX) l = ORC_LOGK_MAX;
  *logk = l;
}

static double log_sinh(double k) { return k + log1p(-exp(-2.0 * k)) - log(2.0); }
static double log_cosh(double z) {
  const double a = fabs(z);
  return a + log1p(exp(-2.0 * a)) - log(2.0);
}

double orc_vmf_logq(const double mu[3], double kappa, const double n[3]) {
  const double c = mu[0] * n[0] + mu[1] * n[1] + mu[2] * n[2];
  return log(kappa) - log(4.0 * M_PI) - log_sinh(kappa) + log_cosh(kappa * c);
}

double orc_mlp_nll_grad(const float *P, int B, const float *X, const float *base, const float *n, int bf16,
                        double *g) {
  memset(g, 0, sizeof(double) * ORC_MLP_NPARAM);
  double loss = 0.0;
  for (int s = 0; s < B; ++s) {
    double xin[22], p1[64], h1[64], p2[64], h2[64], out[4];
    forward_full(P, X + 22 * s, bf16, xin, p1, h1, p2, h2, out);
    const float *bs = base + 3 * s;
    double v[3] = {bs[0] + out[0], bs[1] + out[1], bs[2] + out[2]};
    double r = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (r == 0) r = 1;
    const double mu[3] = {v[0] / r, v[1] / r, v[2] / r};
    double l = out[3];
    int clamped = 0;
    if (l < ORC_LOGK_MIN) {
      l = ORC_LOGK_MIN;
      clamped = 1;
    }
    if (l > ORC_LOGK_MAX) {
      l = ORC_LOGK_MAX;
      clamped = 1;
    }
    const double k = exp(l);
    const double nn[3] = {n[3 * s], n[3 * s + 1], n[3 * s + 2]};
    const double c = mu[0] * nn[0] + mu[1] * nn[1] + mu[2] * nn[2];
    const double z = k * c;
    loss += -l + log(4.0 * M_PI) + log_sinh(k) - log_cosh(z);
    double dout[4];
    const double th = tanh(z);
    dout[3] = clamped ? 0.0 : (-1.0 + k / tanh(k) - z * th);
    double dmu[3] = {-k * th * nn[0], -k * th * nn[1], -k * th * nn[2]};
    const double proj = dmu[0] * mu[0] + dmu[1] * mu[1] + dmu[2] * mu[2];
    for (int a = 0; a < 3; ++a) dout[a] = (dmu[a] - proj * mu[a]) / r;
    /* layer 3 */
    double dh2[64];
    for (int kk = 0; kk < 64; ++kk) dh2[kk] = 0.0;
    for (int o = 0; o < 4; ++o) {
      g[B3_OFF + o] += dout[o];
      for (int kk = 0; kk < 64; ++kk) {
        g[W3_OFF + 64 * o + kk] += dout[o] * h2[kk];
        dh2[kk] += (double)P[W3_OFF + 64 * o + kk] * dout[o];
      }
    }
    for (int j = 0; j < 64; ++j) dh2[j] = p2[j] > 0 ? dh2[j] : 0.0;
    /* layer 2 */
    double dh1[64];
    for (int kk = 0; kk < 64; ++kk) dh1[kk] = 0.0;
    for (int j = 0; j < 64; ++j) {
      g[B2_OFF + j] += dh2[j];
      for (int kk = 0; kk < 64; ++kk) {
        g[W2_OFF + 64 * j + kk] += dh2[j] * h1[kk];
        dh1[kk] += (double)P[W2_OFF + 64 * j + kk] * dh2[j];
      }
    }
    for (int j = 0; j < 64; ++j) dh1[j] = p1[j] > 0 ? dh1[j] : 0.0;
    /* layer 1 */
    for (int j = 0; j < 64; ++j) {
      g[B1_OFF + j] += dh1[j];
      for (int kk = 0; kk < 22; ++kk) g[W1_OFF + 22 * j + kk] += dh1[j] * xin[kk];
    }
  }
  for (int p = 0; p < ORC_MLP_NPARAM; ++p) g[p] /= B;
  return loss / B;
}
