// Two-site Metropolis update shared by the segment kernel (one colour per
// launch) and the slab kernel (both colours per launch).
//
// A thread owns a 16-byte x segment [4q, 4q+4) of one row; the two sites of
// the colour being updated sit at elements PAR and PAR+2.  The caller gathers
// the six neighbours of both sites (bond order +x, -x, +y, -y, +z, -z, the
// order of src/evolve.cpp:119-129) and this header proposes, decides and
// returns the two decisions.
//
// Randomness.  One Philox4x32-10 block per pair (counter = (index of the first
// site, sweep, stream, 0x4D44)): words 0/2 give the angle uniforms, words 1/3
// the accept uniforms (top 24 bits each) and the low bytes of the four words
// the two 16-bit axis uniforms.  The counter depends on the site and the sweep
// only, so a workgroup that recomputes a neighbour's update (the slab kernel's
// halo) reaches the same proposal and decision.
//
// Accept decision.  The energy difference is first formed in FP32: for
// Lebwohl-Lasher with FMAs, sum_b c*((q.m_b)^2 - (o.m_b)^2); for GHRL with
// FMAs from a form of the bond whose constant terms cancel between the two
// directors; for pear bond by bond with the reference's FP64 steps in FP32.
// Its distance
// from the reference's own FP32/FP64 arithmetic has a proven bound (below),
// and the band eps around exp(-dE/T) keeps a several-fold margin over it.  A
// hardware exp2 decides every move whose uniform lies outside the band; moves
// inside it (a few in 10^4 to 10^2 at T ~ 1) are decided with the exact
// reference arithmetic, so every decision equals the one the CPU path makes.
// Surface sites are always decided exactly.
#pragma once
#include "mc_kernels.h"
#include "mc_ghrl_screen.h"

namespace mcsu {

__device__ __forceinline__ float sel(const float4 &v, int e) {
  return e == 0 ? v.x : (e == 1 ? v.y : (e == 2 ? v.z : v.w));
}

// 1 accept, 0 reject, -1 inside the error band (decide exactly).
// y = -dE/T log2(e) is the screen's exp2 argument, formed as acc * ky with
// ky = RN(RN(RN(1/T) log2(e)) c) from the pass constants (c the potential's
// scale of the accumulated bond sum): four roundings in all, as many as the
// former -RN(dE) RN(1/T) log2(e) chain, so y is within 2^-22 |y| of its exact
// value and exp2(y) within ln2 |y| 2^-22 = |x| 2^-22 relative (x = y ln 2 =
// -dE/T), below the |x| 2^-20 term kept in the band; the hardware exp2's own
// error is below the 2^-19 term of relc = RN(eps/T) + 2^-19.  No range check
// is needed: for y above 128 the exp2 is +inf and the band test sends the
// move to the exact path; below -150 it is 0 and u > 0 rejects, as
// expf(x) = 0 does; in between |x| < 104 keeps rel far below 1.
__device__ __forceinline__ int fast_decide(float y, float u, float relc) {
  const float rel = fmaf(fabsf(y), 6.60973460e-7f, relc);  // + |x| 2^-20 = |y| ln2 2^-20
  const float p = __builtin_amdgcn_exp2f(y);
  // Branch free (bitwise |), so the two sites of a pair interleave: a zero
  // uniform is always decided exactly.
  const bool band = (fabsf(u - p) <= p * rel) | (u == 0.0f);
  return band ? -1 : (u < p ? 1 : 0);
}

struct Site {
  float o[3];     // current director
  float q[3];     // proposal
  float m[6][3];  // neighbours +x,-x,+y,-y,+z,-z
  float u;        // accept uniform
};

// Where the pair sits: bit b of `on` set when neighbour b exists (a free z
// face drops one), surf the surface of its plane (-1 for none).  Both sites
// of a pair share one row, so one Where serves both.
struct Where {
  uint32_t on;
  int surf;
};
constexpr Where kBulk{0x3Fu, -1};

// Site energies in the order of the reference's site energy
// (src/geometry.cpp:19-72): bonds summed in FP64 and rounded, then the
// surface term, then the field.  GEO: the lattice may have free z faces and
// surfaces (w matters); without it w is the bulk one and the code is the
// periodic lattice's alone.
template <int POT, bool GEO>
__device__ __forceinline__ void exact_energies(const MCDevParams &P, double cA, float elecE, const Site &S,
                                               const Where &w, float &Eo, float &En) {
  double so = 0.0, sn = 0.0;
#pragma unroll
  for (int b = 0; b < 6; ++b) {
    if (GEO && !((w.on >> b) & 1u)) continue;
    so += (double)mcd_bond<POT>(P, cA, b, S.o[0], S.o[1], S.o[2], S.m[b][0], S.m[b][1], S.m[b][2]);
    sn += (double)mcd_bond<POT>(P, cA, b, S.q[0], S.q[1], S.q[2], S.m[b][0], S.m[b][1], S.m[b][2]);
  }
  Eo = (float)so;
  En = (float)sn;
  if (GEO && w.surf >= 0) {
    const float sz = w.surf ? 1.0f : -1.0f;
    Eo += mcd_surface(P, w.surf, S.o[0], S.o[1], S.o[2], 0.0f, 0.0f, sz);
    En += mcd_surface(P, w.surf, S.q[0], S.q[1], S.q[2], 0.0f, 0.0f, sz);
  }
  if (P.elecA != 0.0f) {
    Eo += mcd_field(P, elecE, S.o[0], S.o[1], S.o[2]);
    En += mcd_field(P, elecE, S.q[0], S.q[1], S.q[2]);
  }
}

// Proposal of src/evolve.cu:143-163 with a 16-bit axis uniform.  The angle
// in revolutions, h = pi*va / (2 pi) = 0.5 (2 u0 - 1) ang_var with
// u0 = (w >> 8) 2^-24, is one fma of the 24-bit integer with the pass
// constants hs = ang_var 2^-24 and hb = -0.5 ang_var: the exact value of
// 0.5 (2 u0 - 1) ang_var rounded once, the bits of 0.5 * ((2 u0 - 1) * av).
__device__ __forceinline__ void propose(const float o[3], float hs, float hb, uint32_t w_angle, uint32_t axis16,
                                        float q[3]) {
  const float h = fmaf((float)(w_angle >> 8), hs, hb);  // sin(pi*va) = sin(2*pi*h)
  const float sn = __builtin_amdgcn_sinf(h), cs = __builtin_amdgcn_cosf(h);
  // Rotation about x (u < 0.333), y (u < 0.666) or z, as one rotation of the
  // pair (a, b) = (o1, o2), (o2, o0) or (o0, o1):
  //   x: q1 = o1 cs + o2 sn, q2 = o2 cs - o1 sn
  //   y: q2 = o2 cs + o0 sn, q0 = o0 cs - o2 sn
  //   z: q0 = o0 cs + o1 sn, q1 = o1 cs - o0 sn
  // Selects instead of branches (the axis differs from lane to lane).  The
  // production proposal need not repeat the reference's roundings (a
  // recorded sweep hands the oracle the proposals it made), so the products
  // are fused.
  const bool ax = axis16 < 21824u, ay = !ax && axis16 < 43647u;
  const float a = ax ? o[1] : (ay ? o[2] : o[0]);
  const float b = ax ? o[2] : (ay ? o[0] : o[1]);
  const float an = fmaf(a, cs, b * sn);
  const float bn = fmaf(b, cs, -(a * sn));
  const float q0 = ax ? o[0] : (ay ? bn : an);
  const float q1 = ax ? an : (ay ? o[1] : bn);
  const float q2 = ax ? bn : (ay ? an : o[2]);
  // Renormalised to first order: |q|^2 = 1 + d with d a few ulps (a rotation
  // of a unit vector), and q (1.5 - 0.5 |q|^2) has |.|^2 = 1 - 0.75 d^2 + O(d^3),
  // closer to 1 than a hardware rsq (and a full-rate fma instead of a
  // transcendental); a director handed in up to 1e-3 off unit length
  // converges quadratically over its accepted moves.
  const float inv = fmaf(-0.5f, fmaf(q2, q2, fmaf(q1, q1, q0 * q0)), 1.5f);
  q[0] = q0 * inv;
  q[1] = q1 * inv;
  q[2] = q2 * inv;
}

// Per-pass constants of the update.
struct PassConst {
  float T, elecE, av;
  uint32_t sweep;
  double cA;
  float cAf, eps;
  float ky1, kyLL;  // -RN(1/T) log2(e) and that times cAf: exp2 arguments of dE and of the LL bond sum
  float relc;       // RN(eps/T) + 2^-19 for the FP32 screen
  McGhrlConst ghrl;  // GHRL screen constants
  float hs, hb;     // ang_var 2^-24 and -0.5 ang_var (propose)
  const uint64_t *etab;  // mcx_expf's table for the exact path (global memory unless a kernel points it at LDS)
};

// Bound on one bond's magnitude from each term's coefficient (unit vectors):
//   GHRL: |E1 (rho P2 + lambda)| <= 2 (|rho| + |lambda|), |mu (a_i a_j n_ij - 1/9)| <= 1.12 |mu|,
//         |nu P2| <= |nu|, |sigma cross| <= |sigma|;
//   pear: |A n^2 + B1 n| <= |A| + |B1|, |0.25 C (1 + 2n + n^2) d| <= 0.25 |C| * 4 * 2.
__device__ __forceinline__ float bond_bound(const MCDevParams &P) {
  if (P.potential == MCD_GHRL)
    return 2.0f * (fabsf(P.rho) + fabsf(P.lambda)) + 1.12f * fabsf(P.mu) + fabsf(P.nu) + fabsf(P.sigma);
  if (P.potential == MCD_PEAR) return fabsf(P.A) + fabsf(P.B1) + 2.0f * fabsf(P.C);
  return 1.5f * fabsf(P.A);
}

__device__ __forceinline__ PassConst pass_const(const MCDevParams &P, const MCDevState *st) {
  PassConst k;
  k.T = st->T;
  k.elecE = st->elecE;
  k.av = st->ang_var;
  k.sweep = st->sweep;
  k.cA = (double)(-P.A) * 1.5;
  k.cAf = (float)k.cA;
  const float e2 = P.eX * P.eX + P.eY * P.eY + P.eZ * P.eZ;
  const float F = fabsf(P.elecA) * k.elecE * k.elecE * e2;
  // 2^-14 * (...): LL with the bound 1.5|A| itself (screen_ll); GHRL and pear
  // with twice the bond bound (screen_ghrl, screen_bonds)
  const float M = bond_bound(P);
  k.eps = 6.1035156e-5f * ((P.potential == MCD_LL ? M : 2.0f * M) + F);
  k.ghrl = mcx_ghrl_const(P.rho, P.lambda, P.mu, P.nu, P.sigma);
  k.ky1 = -((1.0f / k.T) * 1.44269504088896341f);
  k.kyLL = k.cAf * k.ky1;
  k.relc = k.eps / k.T + 1.9073486e-6f;
  k.hs = k.av * 5.96046448e-8f;
  k.hb = -0.5f * k.av;
  k.etab = mcd_exp_table;
  return k;
}

// FP32 screen of one Lebwohl-Lasher move: 1 accept, 0 reject, -1 decide
// exactly.  sum_b (q.m_b)^2 - (o.m_b)^2 = sum_b ((q-o).m_b) ((q+o).m_b).
// Error (u = 2^-24, unit vectors, |q-o| |q+o| <= 2): each dot <= 4u|.|, each
// bond product <= 16u, the running sum <= 72u, so <= 168u * 1.5|A| after the
// scale; the reference's own float/double rounding of the two site energies
// and their difference is <= 114u * 1.5|A|; with the final product, the
// distance to the reference's dE stays below 300u * 1.5|A|, under a third of
// eps = 2^-14 * 1.5|A| = 1024u * 1.5|A|.
// An absent neighbour is a zero vector here, whose bond term is exactly 0;
// a surface site is always decided exactly.
__device__ __forceinline__ float screen_ll_sum(const Site &S) {
  const float d0 = S.q[0] - S.o[0], d1 = S.q[1] - S.o[1], d2 = S.q[2] - S.o[2];
  const float s0 = S.q[0] + S.o[0], s1 = S.q[1] + S.o[1], s2 = S.q[2] + S.o[2];
  float acc = 0.0f;
#pragma unroll
  for (int b = 0; b < 6; ++b) {
    const float e1 = fmaf(d2, S.m[b][2], fmaf(d1, S.m[b][1], d0 * S.m[b][0]));
    const float e2 = fmaf(s2, S.m[b][2], fmaf(s1, S.m[b][1], s0 * S.m[b][0]));
    acc = fmaf(e1, e2, acc);
  }
  return acc;
}

__device__ __forceinline__ float field_delta(const MCDevParams &P, const PassConst &k, const Site &S) {
  return mcd_field(P, k.elecE, S.q[0], S.q[1], S.q[2]) - mcd_field(P, k.elecE, S.o[0], S.o[1], S.o[2]);
}

// FP32 screen of one pear move: each bond of the reference
// (src/potential.cpp:10-15) with every FP64 step done in FP32.  n.r and m.r
// are read off as signed components (r is a lattice axis), the rest is the
// reference's FP32 arithmetic.  Error (u = 2^-24, unit vectors, M =
// bond_bound): the reference's n_ij and the products are the same FP32
// operations here; f - 0.25 C ps is off by <= 3u M and the FP32 additions of
// a bond by <= 2u M.  So each bond is within 12u M of the reference's (FP64
// sum rounded once), twelve of them within 144u M, the FP32 accumulation of
// the twelve differences adds <= 48u M, the reference's own roundings of its
// two site energies and their difference <= 36u M, the field terms (the same
// FP32 operations) <= 4u F: <= 230u M + 4u F against eps = 2^-14 (2M + F) =
// 2048u M + 1024u F, an eightfold margin.
__device__ __forceinline__ float bond32_pear(const MCDevParams &P, int b, const float n[3], const float m[3]) {
  const int a = b >> 1;
  const float sg = (b & 1) ? -1.0f : 1.0f;
  const float nij = mcd_dot(n[0], n[1], n[2], m[0], m[1], m[2]);
  float aa = 1.0f + 2.0f * nij;
  aa = aa + nij * nij;
  const float d = sg * (m[a] - n[a]);
  const float ps = aa * d;
  float f = -P.A * nij * nij;
  f = f - P.B1 * nij;
  return f - (P.C * 0.25f) * ps;
}

template <bool GEO>
__device__ __forceinline__ float screen_pear_sum(const MCDevParams &P, const Site &S, const Where &w) {
  float acc = 0.0f;
#pragma unroll
  for (int b = 0; b < 6; ++b) {
    if (GEO && !((w.on >> b) & 1u)) continue;
    acc += bond32_pear(P, b, S.q, S.m[b]) - bond32_pear(P, b, S.o, S.m[b]);
  }
  return acc;
}

// Screens both sites of a pair (d = 1 accept, 0 reject, -1 decide exactly).
// Their arithmetic is straight-line code for both sites, so the compiler
// interleaves the two independent chains; the field is one branch on a
// uniform value for both.  Surface sites are always decided exactly.
//   LL:   y = acc kyLL (+ fd ky1 with the field), acc the bond sum above.
//   GHRL: dE = mcx_ghrl_delta (mc_ghrl_screen.h), within 920u M (+ 6u F with
//         the field) of the reference's dE against eps = 2^-14 (2M + F) =
//         2048u M + 1024u F, over twice that; a tiny denominator there sends
//         the site to the exact path.
//   pear: dE = the FP32 bond sum above (+ fd).
template <int POT, bool GEO>
__device__ __forceinline__ void screen_pair(const MCDevParams &P, const PassConst &k, const Site S[2], const Where &w,
                                            int &d0, int &d1) {
  if (GEO && w.surf >= 0) {
    d0 = d1 = -1;
    return;
  }
  float y0, y1;
  bool ex0 = false, ex1 = false;
  if (POT == MCD_LL) {
    const float a0 = screen_ll_sum(S[0]), a1 = screen_ll_sum(S[1]);
    if (P.elecA != 0.0f) {
      y0 = fmaf(a0, k.kyLL, field_delta(P, k, S[0]) * k.ky1);
      y1 = fmaf(a1, k.kyLL, field_delta(P, k, S[1]) * k.ky1);
    } else {
      y0 = a0 * k.kyLL;
      y1 = a1 * k.kyLL;
    }
  } else {
    float e0, e1;
    if (POT == MCD_GHRL) {
      e0 = mcx_ghrl_delta(k.ghrl, S[0].o, S[0].q, S[0].m, GEO ? w.on : 0x3Fu, &ex0);
      e1 = mcx_ghrl_delta(k.ghrl, S[1].o, S[1].q, S[1].m, GEO ? w.on : 0x3Fu, &ex1);
    } else {
      e0 = screen_pear_sum<GEO>(P, S[0], w);
      e1 = screen_pear_sum<GEO>(P, S[1], w);
    }
    if (P.elecA != 0.0f) {
      e0 += field_delta(P, k, S[0]);
      e1 += field_delta(P, k, S[1]);
    }
    y0 = e0 * k.ky1;
    y1 = e1 * k.ky1;
  }
  d0 = fast_decide(y0, S[0].u, k.relc);
  d1 = fast_decide(y1, S[1].u, k.relc);
  if (ex0) d0 = -1;
  if (ex1) d1 = -1;
}

// The neighbours of the pair: yp..zm hold, per component, the +y/-y/+z/-z
// neighbour of site 0 in .x and of site 1 in .y; ov is the x neighbour outside
// the segment (x = 4q-1 for PAR 0, x = 4q+4 for PAR 1).
struct PairNbrs {
  float2 yp[3], ym[3], zp[3], zm[3];
  float ov[3];
};

// Proposes and decides both sites of the pair whose first site has linear
// index sA; c is the segment.  Fills S (for recording) and a.
// An absent z neighbour (free face) must arrive as a zero vector with its
// bit clear in w.on.
template <int POT, int PAR, bool GEO>
__device__ __forceinline__ void pair_update(const MCDevParams &P, const PassConst &k, const float4 c[3],
                                            const PairNbrs &nb, uint32_t sA, const Where &w, Site S[2], bool a[2],
                                            MCU4 &r) {
  r = mcd_philox(MCU4{sA, k.sweep, P.stream, 0x4D44u}, P.seed_lo, P.seed_hi);
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const int e = PAR + 2 * t;  // element of this site in the segment
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      S[t].o[d] = sel(c[d], e);
      S[t].m[0][d] = (e == 3) ? nb.ov[d] : sel(c[d], e + 1);  // +x
      S[t].m[1][d] = (e == 0) ? nb.ov[d] : sel(c[d], e - 1);  // -x
      S[t].m[2][d] = t ? nb.yp[d].y : nb.yp[d].x;
      S[t].m[3][d] = t ? nb.ym[d].y : nb.ym[d].x;
      S[t].m[4][d] = t ? nb.zp[d].y : nb.zp[d].x;
      S[t].m[5][d] = t ? nb.zm[d].y : nb.zm[d].x;
    }
  }
  const uint32_t ax0 = ((r.x & 0xFFu) << 8) | (r.y & 0xFFu);
  const uint32_t ax1 = ((r.z & 0xFFu) << 8) | (r.w & 0xFFu);
  propose(S[0].o, k.hs, k.hb, r.x, ax0, S[0].q);
  propose(S[1].o, k.hs, k.hb, r.z, ax1, S[1].q);
  S[0].u = mcd_u01(r.y);
  S[1].u = mcd_u01(r.w);
  // Screen both sites, then one shared (rarely executed) exact evaluation for
  // whichever sites fell inside the error band.
  int d0, d1;
  screen_pair<POT, GEO>(P, k, S, w, d0, d1);
  a[0] = d0 == 1;
  a[1] = d1 == 1;
  if (d0 < 0 || d1 < 0) {
#pragma unroll 1
    for (int t = 0; t < 2; ++t) {
      if ((t ? d1 : d0) >= 0) continue;
      const Site &X = t ? S[1] : S[0];
      float Eo, En;
      exact_energies<POT, GEO>(P, k.cA, k.elecE, X, w, Eo, En);
      const bool at = mcd_accept_with(En, Eo, k.T, X.u, k.etab);
      if (t) a[1] = at; else a[0] = at;
    }
  }
}

// Writes the accepted proposals into the segment.
template <int PAR>
__device__ __forceinline__ void pair_apply(float4 c[3], const Site S[2], const bool a[2]) {
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (a[0]) {
      if (PAR == 0) c[d].x = S[0].q[d]; else c[d].y = S[0].q[d];
    }
    if (a[1]) {
      if (PAR == 0) c[d].z = S[1].q[d]; else c[d].w = S[1].q[d];
    }
  }
}

// Record of both sites of a pair (the parity replay reads it back).
template <int POT, bool GEO>
__device__ __forceinline__ void pair_record(const MCDevParams &P, const PassConst &k, const MCRecord &rec, int sA,
                                            const Where &w, const Site S[2], const bool a[2], const MCU4 &r) {
  const uint32_t ax0 = ((r.x & 0xFFu) << 8) | (r.y & 0xFFu);
  const uint32_t ax1 = ((r.z & 0xFFu) << 8) | (r.w & 0xFFu);
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const int s = sA + 2 * t;
    float Eo, En;
    exact_energies<POT, GEO>(P, k.cA, k.elecE, S[t], w, Eo, En);
    rec.u0[s] = mcd_u01(t ? r.z : r.x);
    rec.u1[s] = (float)(t ? ax1 : ax0) * (1.0f / 65536.0f);
    rec.u2[s] = S[t].u;
    rec.px[s] = S[t].q[0];
    rec.py[s] = S[t].q[1];
    rec.pz[s] = S[t].q[2];
    rec.acc[s] = a[t] ? 1 : 0;
    rec.eo[s] = Eo;
    rec.en[s] = En;
  }
}

}  // namespace mcsu
