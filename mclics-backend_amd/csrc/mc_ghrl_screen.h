// The FP32 energy difference the GHRL screen forms for one move
// (src/potential.cpp:22-38), shared by the sweep kernels and a host check of
// its error bound.
//
// For bond b along axis a with sign s (r = s e_a), x = n.m, P2 = 1.5 x^2 - 0.5:
//   E = E1 (rho P2 + lambda) + mu (n_a m_a x - 1/9) + nu P2 + sigma sgn(x) s (m x n)_a,
//   E1 = 1.5 n_a^2 + 1.5 m_a^2 - 1,
// so rho P2 + lambda = 1.5 rho x^2 + (lambda - 0.5 rho), and the constants
// -mu/9 and -0.5 nu cancel between the proposal's bond and the current one.
// Each bond difference is formed as bq - bo with
//   b = (1.5 n_a^2 + (1.5 m_a^2 - 1)) (1.5 rho x^2 + lambda - 0.5 rho) + mu (n_a m_a) x + 1.5 nu x^2
// plus the chiral part sigma s (sgn(x_q)(m x q)_a - sgn(x_o)(m x o)_a) when
// sigma != 0, with fused multiply-adds throughout.
//
// Error (u = 2^-24, unit vectors, M = 2(|rho|+|lambda|) + 1.12|mu| + |nu| +
// |sigma|): x is within 3u of n.m, and of the reference's n_ij within 6u.
// The chiral sign cannot differ from the reference's unless |x| <= 6u; such
// a move (|x| < 2^-20 on a chiral lattice) sets *tiny and is decided
// exactly.  Per evaluation of b: the E1 (..) term is within 38u|rho| +
// 14u|lambda| of its exact value (x^2 within 7u, the constants rounded once,
// E1 within 8u), the mu term within 5u|mu|, the nu term within 12u|nu|, the
// cross product within 2u|sigma|: <= 38u M.  Twelve evaluations: 456u M.
// The roundings of the accumulation (partial sums of b within 1.5M, of a
// bond difference within 5M, of the site sum within 30M): <= 54u M + 48u M
// + 180u M.  The reference's own roundings, bond by bond (12u M each) and of
// its site energies and their difference (36u M): 180u M.  In all <= 920u M
// from the reference's dE (before the field, whose FP32 values are the
// reference's own: 6u F more).
#pragma once
#include <math.h>
#include <stdint.h>

#ifndef MCX_HD
#define MCX_HD __host__ __device__
#endif

// Screen constants: RN(1.5 rho), RN(lambda - 0.5 rho), RN(1.5 nu).
struct McGhrlConst {
  float g15r, glr, g15n, mu, sigma;
};

MCX_HD inline McGhrlConst mcx_ghrl_const(float rho, float lambda, float mu, float nu, float sigma) {
  McGhrlConst c;
  c.g15r = 1.5f * rho;
  c.glr = lambda - 0.5f * rho;
  c.g15n = 1.5f * nu;
  c.mu = mu;
  c.sigma = sigma;
  return c;
}

// dE of moving o to q with neighbours m[b] (bond order +x,-x,+y,-y,+z,-z;
// bonds whose bit in `on` is clear are left out).
MCX_HD inline float mcx_ghrl_delta(const McGhrlConst &c, const float o[3], const float q[3], const float m[6][3],
                                   uint32_t on, bool *tiny) {
  const bool chiral = c.sigma != 0.0f;
  float hq[3], ho[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    hq[a] = 1.5f * q[a] * q[a];
    ho[a] = 1.5f * o[a] * o[a];
  }
  float acc = 0.0f;
  bool small = false;
#pragma unroll
  for (int b = 0; b < 6; ++b) {
    if (!((on >> b) & 1u)) continue;
    const int a = b >> 1, a1 = (a + 1) % 3, a2 = (a + 2) % 3;
    const float *mb = m[b];
    const float cm = fmaf(1.5f * mb[a], mb[a], -1.0f);
    const float xq = fmaf(q[2], mb[2], fmaf(q[1], mb[1], q[0] * mb[0]));
    const float xo = fmaf(o[2], mb[2], fmaf(o[1], mb[1], o[0] * mb[0]));
    const float xq2 = xq * xq, xo2 = xo * xo;
    float bq = (hq[a] + cm) * fmaf(c.g15r, xq2, c.glr);
    float bo = (ho[a] + cm) * fmaf(c.g15r, xo2, c.glr);
    bq = fmaf(c.mu, (q[a] * mb[a]) * xq, bq);
    bo = fmaf(c.mu, (o[a] * mb[a]) * xo, bo);
    bq = fmaf(c.g15n, xq2, bq);
    bo = fmaf(c.g15n, xo2, bo);
    float db = bq - bo;
    if (chiral) {
      // the reference's cross product: (m x n)_a = n_{a+2} m_{a+1} - n_{a+1} m_{a+2}
      const float cq = fmaf(q[a2], mb[a1], -(q[a1] * mb[a2]));
      const float co = fmaf(o[a2], mb[a1], -(o[a1] * mb[a2]));
      const float sq = xq > 0.0f ? cq : -cq, so = xo > 0.0f ? co : -co;
      db = fmaf((b & 1) ? -c.sigma : c.sigma, sq - so, db);
      small = small || fabsf(xq) < 9.53674316e-7f || fabsf(xo) < 9.53674316e-7f;
    }
    acc += db;
  }
  *tiny = small;
  return acc;
}
