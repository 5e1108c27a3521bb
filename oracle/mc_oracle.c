/*
 * CPU oracle for the McLiCS checkerboard Metropolis sweep (see mc_oracle.h).
 *
 * TEST INFRASTRUCTURE ONLY: a C restatement of the reference CPU path used as
 * the checker.  Built with `gcc -O3 -fopenmp` and no -march (reference
 * Makefile:13 minus -static), so FP32 expressions are evaluated one IEEE
 * operation at a time without FMA contraction, exactly as the reference's g++
 * build evaluates them.
 */
#include "mc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "philox.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

static inline float dot3f(const float a[3], const float b[3]) {
  /* ((a0*b0 + a1*b1) + a2*b2) in FP32, as written in src/potential.cpp:18. */
  float t = a[0] * b[0];
  t = t + a[1] * b[1];
  t = t + a[2] * b[2];
  return t;
}

/* src/potential.cpp:17-20: -params->A * 1.5 * nij * nij, the 1.5 a double
 * literal, so the chain is evaluated in double and rounded to float once. */
static float bond_ll(const orc_params *p, const float ni[3], const float nj[3]) {
  float nij = dot3f(ni, nj);
  double v = (double)(-p->A) * 1.5;
  v = v * (double)nij;
  v = v * (double)nij;
  return (float)v;
}

/* src/potential.cpp:22-38.  Types: E1 is a float holding a double
 * expression; em*(ai*aj*nij - 1.0/9) is double; the whole sum is double and
 * rounded to float on return. */
static float bond_ghrl(const orc_params *p, const float ni[3], const float nj[3], const float r[3]) {
  const float el = p->ghrl_lambda, em = p->ghrl_mu, en = p->ghrl_nu, er = p->ghrl_rho, es = p->ghrl_sigma;
  float ai = dot3f(ni, r);
  float aj = dot3f(nj, r);
  float nij = dot3f(ni, nj);
  float pij = 1.5f * nij * nij - 0.5f;
  float c0 = ni[2] * nj[1] - ni[1] * nj[2];
  float c1 = ni[0] * nj[2] - ni[2] * nj[0];
  float c2 = ni[1] * nj[0] - ni[0] * nj[1];
  float cross = c0 * r[0];
  cross = cross + c1 * r[1];
  cross = cross + c2 * r[2];
  double e1d = 1.5 * (double)ai * (double)ai;
  e1d = e1d + 1.5 * (double)aj * (double)aj;
  e1d = e1d - 1.0;
  float E1 = (float)e1d;
  float t1 = E1 * (er * pij + el);
  float aan = ai * aj * nij;
  double t2 = (double)em * ((double)aan - 1.0 / 9);
  double s = (double)t1 + t2;
  s = s + (double)(en * pij);
  float sg = (float)(nij > 0 ? 1 : -1);
  s = s + (double)(es * sg * cross);
  return (float)s;
}

/* src/potential.cpp:10-15.  (C*0.25) is double, so the last subtraction is
 * done in double. */
static float bond_pear(const orc_params *p, const float ni[3], const float nj[3], const float r[3]) {
  float nij = dot3f(ni, nj);
  float a = 1.0f + 2.0f * nij;
  a = a + nij * nij;
  float d = (nj[0] - ni[0]) * r[0];
  d = d + (nj[1] - ni[1]) * r[1];
  d = d + (nj[2] - ni[2]) * r[2];
  float ps = a * d;
  float f = -p->A * nij * nij;
  f = f - p->B1 * nij;
  double v = (double)f - ((double)p->C * 0.25) * (double)ps;
  return (float)v;
}

float orc_bond(const orc_params *p, const float ni[3], const float nj[3], const float r[3]) {
  switch (p->potential) {
    case ORC_GHRL:
      return bond_ghrl(p, ni, nj, r);
    case ORC_PEAR:
      return bond_pear(p, ni, nj, r);
    default:
      return bond_ll(p, ni, nj);
  }
}

/* src/potential.cpp:39-43. */
float orc_field(const orc_params *p, const float ni[3]) {
  float d = ni[0] * p->elecX;
  d = d + ni[1] * p->elecY;
  d = d + ni[2] * p->elecZ;
  float nde = p->elecE * d;
  return -p->elecA * (nde * nde);
}

/* Shared GHRL surface form of src/anchoring_homeotropic.cpp:51-68,
 * src/anchoring_rp.cpp:77-98 and src/anchoring_fg.cpp:51-74, all in FP32;
 * the (1/9) there is integer division and contributes 0. */
static float ghrl_surface_form(const orc_params *p, float W, const float ni[3], const float nj[3],
                               const float s[3], float ai, float aj) {
  const float el = p->ghrl_lambda, em = p->ghrl_mu, en = p->ghrl_nu, er = p->ghrl_rho, es = p->ghrl_sigma;
  float nij = dot3f(ni, nj);
  float pij = 1.5f * nij * nij - 0.5f;
  float c0 = ni[2] * nj[1] - ni[1] * nj[2];
  float c1 = ni[0] * nj[2] - ni[2] * nj[0];
  float c2 = ni[1] * nj[0] - ni[0] * nj[1];
  float cross = c0 * s[0];
  cross = cross + c1 * s[1];
  cross = cross + c2 * s[2];
  float E1 = 1.5f * ai * ai;
  E1 = E1 + 1.5f * aj * aj;
  E1 = E1 - 1.0f;
  float t = E1 * (er * pij + el);
  t = t + em * (ai * aj * nij);
  t = t - 0.0f;
  t = t + en * pij;
  float sg = (float)(nij > 0 ? 1 : -1);
  t = t + es * sg * cross;
  return W * t;
}

float orc_surface(const orc_params *p, int surf, const float ni[3], const float s[3]) {
  const float W = p->surf_W[surf];
  const float *d = p->surf_dir[surf];
  switch (p->surf_kind[surf]) {
    case ORC_HOMEOTROPIC: { /* src/anchoring_homeotropic.cpp:29-34 */
      float nij = dot3f(ni, s);
      return -W * nij * nij;
    }
    case ORC_RP:     /* src/anchoring_rp.cpp:41-48 */
    case ORC_STRONG: /* src/anchoring_strong.cpp:41-48 */
    case ORC_STRONG_GHRL: { /* src/anchoring_strong.cpp:77-84 */
      float nij = dot3f(ni, d);
      return -W * nij * nij;
    }
    case ORC_FG: { /* src/anchoring_fg.cpp:29-34 */
      float nij = dot3f(ni, s);
      return +W * nij * nij;
    }
    case ORC_HOMEOTROPIC_GHRL: {
      float ai = dot3f(ni, s);
      return ghrl_surface_form(p, W, ni, s, s, ai, 1.0f);
    }
    case ORC_RP_GHRL: {
      float ai = dot3f(ni, s);
      float aj = dot3f(d, s);
      return ghrl_surface_form(p, W, ni, d, s, ai, aj);
    }
    case ORC_FG_GHRL: { /* src/anchoring_fg.cpp:59-64: projected partner */
      float mod = sqrtf(fabsf(dot3f(ni, s)));
      float nj[3] = {ni[0] - s[0] * mod, ni[1] - s[1] * mod, ni[2] - s[2] * mod};
      mod = sqrtf(fabsf(dot3f(nj, nj)));
      float dv = mod > 0 ? mod : 1.0f;
      nj[0] /= dv;
      nj[1] /= dv;
      nj[2] /= dv;
      float ai = dot3f(ni, s);
      float aj = dot3f(nj, s);
      return ghrl_surface_form(p, W, ni, nj, s, ai, aj);
    }
    default:
      return 0.0f;
  }
}

/* src/io.cpp:277-302. */
void orc_set_ghrl(orc_params *p, float k11, float k22, float k33, float p0) {
  float Scale = (float)fabs(9.0 / (double)(k11 - 3 * k22 - k33));
  p->ghrl_sigma = p0 != 0 ? (float)(-Scale * k22 * (2 * M_PI / p0)) : 0.0f;
  p->ghrl_lambda = (Scale / 9) * (2 * k11 - 3 * k22 + k33);
  p->ghrl_mu = Scale * (k22 - k11);
  p->ghrl_rho = (Scale / 9) * (k11 - k33);
  p->ghrl_nu = (Scale / 9) * (k11 - 3 * k22 - k33);
}

/* Boundary resolution, src/potential.cpp:44-54.  Returns 1 and wraps *ii
 * for an existing neighbour, 0 when a free boundary drops it. */
static inline int resolve(int bound, int *ii, int N) {
  if (bound == ORC_PERIODIC) {
    *ii = (*ii + N) % N;
    return 1;
  }
  return !(*ii == -1 || *ii == N);
}

/* src/geometry.cpp:19-72 plus src/geometry_{bulk,slab}.cpp latice_Potential. */
float orc_site_energy(const orc_params *p, const float *n, const int32_t *pt, const float *ns, int i, int j,
                      int k, const float *n0) {
  const int Nx = p->nx, Ny = p->ny, Nz = p->nz;
  const int64_t s = (int64_t)i + (int64_t)Nx * (j + (int64_t)Ny * k);
  float ni[3];
  if (n0) {
    ni[0] = n0[0];
    ni[1] = n0[1];
    ni[2] = n0[2];
  } else {
    ni[0] = n[3 * s + 0];
    ni[1] = n[3 * s + 1];
    ni[2] = n[3 * s + 2];
  }
  static const int off[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
  double E = 0;
  for (int b = 0; b < 6; ++b) {
    int ii = i + off[b][0], jj = j + off[b][1], kk = k + off[b][2];
    int ok;
    if (b < 2)
      ok = resolve(p->bound[0], &ii, Nx);
    else if (b < 4)
      ok = resolve(p->bound[1], &jj, Ny);
    else
      ok = resolve(p->bound[2], &kk, Nz);
    if (!ok) continue;
    int64_t t = (int64_t)ii + (int64_t)Nx * (jj + (int64_t)Ny * kk);
    if (!pt[t]) continue;
    float nj[3] = {n[3 * t + 0], n[3 * t + 1], n[3 * t + 2]};
    float r[3] = {(float)off[b][0], (float)off[b][1], (float)off[b][2]};
    E += orc_bond(p, ni, nj, r);
  }
  float Ef = (float)E;
  if (pt[s] > 1 && p->n_surf > 0) {
    float sv[3] = {ns[3 * s + 0], ns[3 * s + 1], ns[3 * s + 2]};
    Ef += orc_surface(p, pt[s] - 2, ni, sv);
  }
  if (p->elecA != 0) Ef += orc_field(p, ni);
  return Ef;
}

void orc_site_energies(const orc_params *p, const float *n, const int32_t *pt, const float *ns, float *out) {
  const int Nx = p->nx, Ny = p->ny, Nz = p->nz;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < Nz; ++k)
    for (int j = 0; j < Ny; ++j)
      for (int i = 0; i < Nx; ++i) {
        int64_t s = (int64_t)i + (int64_t)Nx * (j + (int64_t)Ny * k);
        out[s] = orc_site_energy(p, n, pt, ns, i, j, k, NULL);
      }
}

/* exp of a float argument as the reference's accept test evaluates it
 * (src/evolve.cpp:160: exp() of a float in C++ is std::exp(float), glibc's
 * expf), over an array; the tests use it to find the inputs where glibc's
 * expf differs from the correctly rounded value. */
void orc_expf_n(const float *x, float *y, int64_t n) {
  for (int64_t i = 0; i < n; ++i) y[i] = expf(x[i]);
}

/* Accept/reject with the reference's semantics (src/evolve.cpp:154-165). */
static inline int update_site(const orc_params *p, float *n, const int32_t *pt, const float *ns, int i, int j,
                              int k, float T, const float nn[3], float u, float *eo, float *en) {
  int64_t s = (int64_t)i + (int64_t)p->nx * (j + (int64_t)p->ny * k);
  float E_old = orc_site_energy(p, n, pt, ns, i, j, k, NULL);
  float E_new = orc_site_energy(p, n, pt, ns, i, j, k, nn);
  if (eo) *eo = E_old;
  if (en) *en = E_new;
  float x = -(E_new - E_old) / T;
  if ((double)u < (double)expf(x)) {
    n[3 * s + 0] = nn[0];
    n[3 * s + 1] = nn[1];
    n[3 * s + 2] = nn[2];
    return 1;
  }
  return 0;
}

int64_t orc_pass_inject(const orc_params *p, float *n, const int32_t *pt, const float *ns, int set, float T,
                        const float *prop, const float *u, uint8_t *acc, float *e_old, float *e_new) {
  const int Nx = p->nx, Ny = p->ny, Nz = p->nz;
  int64_t accepted = 0;
  if (set < 8) {
    /* src/evolve.cpp:92-113: cells in (iBox fastest) order. */
    const int di = set % 2, dj = (set / 2) % 2, dk = set / 4;
    const int Ni = Nx / 2 + Nx % 2, Nj = Ny / 2 + Ny % 2, Nk = Nz / 2 + Nz % 2;
    const int64_t Ntt = (int64_t)Ni * Nj * Nk;
    for (int64_t nt = 0; nt < Ntt; ++nt) {
      int iBox = (int)(nt % Ni), jBox = (int)((nt / Ni) % Nj), kBox = (int)(nt / Ni / Nj);
      int i = di + 2 * iBox, j = dj + 2 * jBox, k = dk + 2 * kBox;
      if (i >= Nx || j >= Ny || k >= Nz) continue;
      int64_t s = (int64_t)i + (int64_t)Nx * (j + (int64_t)Ny * k);
      if (pt[s] == 0) continue;
      float eo, en;
      int a = update_site(p, n, pt, ns, i, j, k, T, prop + 3 * s, u[s], &eo, &en);
      if (acc) acc[s] = (uint8_t)a;
      if (e_old) e_old[s] = eo;
      if (e_new) e_new[s] = en;
      accepted += a;
    }
  } else {
    const int c = set - 8;
    for (int k = 0; k < Nz; ++k)
      for (int j = 0; j < Ny; ++j)
        for (int i = 0; i < Nx; ++i) {
          if (((i + j + k) & 1) != c) continue;
          int64_t s = (int64_t)i + (int64_t)Nx * (j + (int64_t)Ny * k);
          if (pt[s] == 0) continue;
          float eo, en;
          int a = update_site(p, n, pt, ns, i, j, k, T, prop + 3 * s, u[s], &eo, &en);
          if (acc) acc[s] = (uint8_t)a;
          if (e_old) e_old[s] = eo;
          if (e_new) e_new[s] = en;
          accepted += a;
        }
  }
  return accepted;
}

/* src/evolve.cpp:169-178 with float ang_var and double literals. */
float orc_adapt_ang_var(float ang_var, int64_t accepted, int64_t valid) {
  double rate = 1.0 * (double)accepted / (double)valid;
  if (rate < 0.5) {
    ang_var = (float)((double)ang_var * 0.99);
    if (ang_var < 0.01) ang_var = 0.01f;
  } else if (rate > 0.5) {
    ang_var = (float)((double)ang_var / 0.99);
    if (ang_var > 1.0) ang_var = (float)((double)ang_var - 0.5);
  }
  return ang_var;
}

/* Proposal of src/evolve.cpp:136-152 driven by three 24-bit uniforms. */
static inline void propose_cpu(const float ni[3], float ang_var, float u0, float u1, float out[3]) {
  float va = (float)((2.0 * (double)u0 - 1.0) * (double)ang_var);
  float ang = (float)(M_PI * (double)va);
  float c = cosf(ang), s = sinf(ang);
  if ((double)u1 < 0.333) {
    out[0] = ni[0];
    out[1] = ni[1] * c + ni[2] * s;
    out[2] = ni[2] * c - ni[1] * s;
  } else if ((double)u1 < 0.666) {
    out[0] = ni[0] * c - ni[2] * s;
    out[1] = ni[1];
    out[2] = ni[2] * c + ni[0] * s;
  } else {
    out[0] = ni[0] * c + ni[1] * s;
    out[1] = ni[1] * c - ni[0] * s;
    out[2] = ni[2];
  }
}

int64_t orc_sweep_rng(const orc_params *p, float *n, const int32_t *pt, const float *ns, float T, float *ang_var,
                      uint32_t seed_lo, uint32_t seed_hi, uint32_t sweep, uint32_t stream, int order, int nthreads) {
  const int Nx = p->nx, Ny = p->ny, Nz = p->nz;
  const int64_t Nt = (int64_t)Nx * Ny * Nz;
  const float av = *ang_var;
  int64_t accepted = 0, valid = 0;
  for (int64_t s = 0; s < Nt; ++s) valid += pt[s] != 0;
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
  nthreads = 1;
#endif
  const int nsets = order == 0 ? 8 : 2;
#pragma omp parallel num_threads(nthreads) reduction(+ : accepted)
  for (int set = 0; set < nsets; ++set) {
    const int di = set % 2, dj = (set / 2) % 2, dk = set / 4;
    const int Ni = Nx / 2 + Nx % 2, Nj = Ny / 2 + Ny % 2, Nk = Nz / 2 + Nz % 2;
    const int64_t Ntt = order == 0 ? (int64_t)Ni * Nj * Nk : Nt;
#pragma omp for schedule(static)
    for (int64_t nt = 0; nt < Ntt; ++nt) {
      int i, j, k;
      if (order == 0) {
        i = di + 2 * (int)(nt % Ni);
        j = dj + 2 * (int)((nt / Ni) % Nj);
        k = dk + 2 * (int)(nt / Ni / Nj);
        if (i >= Nx || j >= Ny || k >= Nz) continue;
      } else {
        i = (int)(nt % Nx);
        j = (int)((nt / Nx) % Ny);
        k = (int)(nt / ((int64_t)Nx * Ny));
        if (((i + j + k) & 1) != set) continue;
      }
      int64_t s = (int64_t)i + (int64_t)Nx * (j + (int64_t)Ny * k);
      if (pt[s] == 0) continue;
      mc_u32x4 ctr = {{(uint32_t)s, sweep, stream, 0x4D43u}};
      mc_u32x4 r = mc_philox4x32_10(ctr, seed_lo, seed_hi);
      float nn[3];
      propose_cpu(n + 3 * s, av, mc_u01_24(r.v[0]), mc_u01_24(r.v[1]), nn);
      accepted += update_site(p, n, pt, ns, i, j, k, T, nn, mc_u01_24(r.v[2]), NULL, NULL);
    }
  }
  *ang_var = orc_adapt_ang_var(av, accepted, valid);
  return accepted;
}

double orc_energy_mean(const orc_params *p, const float *n, const int32_t *pt, const float *ns, int gpu_semantics,
                       int nthreads) {
  const int Nx = p->nx, Ny = p->ny, Nz = p->nz;
  const int64_t Nt = (int64_t)Nx * Ny * Nz;
  double Etot = 0;
  int64_t valid = 0;
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
  nthreads = 1;
#endif
#pragma omp parallel for num_threads(nthreads) reduction(+ : Etot, valid) schedule(static)
  for (int64_t idx = 0; idx < Nt; ++idx) {
    int i = (int)(idx % Nx);
    int j = (int)((idx / Nx) % Ny);
    int k = (int)(idx / ((int64_t)Nx * Ny));
    if (gpu_semantics && pt[idx] == 0) continue;
    valid += pt[idx] != 0;
    Etot += orc_site_energy(p, n, pt, ns, i, j, k, NULL);
  }
  return gpu_semantics ? Etot / (double)valid : Etot / (double)Nt;
}

void orc_qtensor(const float *n, const int32_t *pt, int64_t nt, double q6[6]) {
  double a[6] = {0, 0, 0, 0, 0, 0};
  int64_t pts = 0;
  for (int64_t s = 0; s < nt; ++s) {
    if (!pt[s]) continue;
    double x = n[3 * s], y = n[3 * s + 1], z = n[3 * s + 2];
    a[0] += 3.0 * x * x - 1.0;
    a[1] += 3.0 * x * y;
    a[2] += 3.0 * x * z;
    a[3] += 3.0 * y * y - 1.0;
    a[4] += 3.0 * y * z;
    a[5] += 3.0 * z * z - 1.0;
    ++pts;
  }
  for (int c = 0; c < 6; ++c) q6[c] = pts ? a[c] / (2.0 * (double)pts) : 0.0;
}

/* Symmetric 3x3 eigenvalues by the trigonometric closed form (Smith, CACM
 * 4(4) 1961), sorted descending. */
void orc_sym3_eigvals(const double q[6], double ev[3]) {
  const double a = q[0], b = q[1], c = q[2], d = q[3], e = q[4], f = q[5];
  double p1 = b * b + c * c + e * e;
  double tr = (a + d + f) / 3.0;
  double l0, l1, l2;
  double p2 = (a - tr) * (a - tr) + (d - tr) * (d - tr) + (f - tr) * (f - tr) + 2.0 * p1;
  if (p2 <= 1e-300) {
    l0 = l1 = l2 = tr;
  } else {
    double pp = sqrt(p2 / 6.0);
    double B0 = (a - tr) / pp, B1 = b / pp, B2 = c / pp, B3 = (d - tr) / pp, B4 = e / pp, B5 = (f - tr) / pp;
    double det = B0 * (B3 * B5 - B4 * B4) - B1 * (B1 * B5 - B4 * B2) + B2 * (B1 * B4 - B3 * B2);
    double r = det / 2.0;
    if (r <= -1.0)
      r = -1.0;
    else if (r >= 1.0)
      r = 1.0;
    double phi = acos(r) / 3.0;
    l0 = tr + 2.0 * pp * cos(phi);
    l2 = tr + 2.0 * pp * cos(phi + 2.0 * M_PI / 3.0);
    l1 = 3.0 * tr - l0 - l2;
  }
  double v[3] = {l0, l1, l2};
  for (int x = 0; x < 3; ++x)
    for (int y = x + 1; y < 3; ++y)
      if (v[y] > v[x]) {
        double t = v[x];
        v[x] = v[y];
        v[y] = t;
      }
  ev[0] = v[0];
  ev[1] = v[1];
  ev[2] = v[2];
}

double orc_local_S(const orc_params *p, const float *n, const int32_t *pt, int i, int j, int k) {
  const int Nx = p->nx, Ny = p->ny, Nz = p->nz;
  double Q[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  int points = 0;
  for (int di = -1; di < 2; ++di)
    for (int dj = -1; dj < 2; ++dj)
      for (int dk = -1; dk < 2; ++dk) {
        if (abs(di) + abs(dj) + abs(dk) > 1) continue;
        int ii = i + di, jj = j + dj, kk = k + dk;
        if (!resolve(p->bound[0], &ii, Nx)) continue;
        if (!resolve(p->bound[1], &jj, Ny)) continue;
        if (!resolve(p->bound[2], &kk, Nz)) continue;
        int64_t t = (int64_t)ii + (int64_t)Nx * (jj + (int64_t)Ny * kk);
        if (!pt[t]) continue;
        for (int l = 0; l < 3; ++l) {
          for (int m = 0; m <= l; ++m) Q[3 * m + l] += 3.0 * ((double)n[3 * t + l] * (double)n[3 * t + m]);
          Q[4 * l] -= 1.0;
        }
        ++points;
      }
  double q6[6] = {Q[0], Q[1], Q[2], Q[4], Q[5], Q[8]};
  for (int c = 0; c < 6; ++c) q6[c] /= 2.0 * points;
  double ev[3];
  orc_sym3_eigvals(q6, ev);
  return ev[0];
}

int orc_float_loop(float xi, float xf, float dx, float *xs, int max_rows) {
  int sign = (int)(-dx / fabsf(dx));
  int rows = 0;
  for (float x = xi; (float)(1000000 * sign) * (x - xf) >= 0; x += dx) {
    if (rows < max_rows && xs) xs[rows] = x;
    ++rows;
    if (rows > 100000000) break;
  }
  return rows;
}

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  mc_u32x4 c = {{ctr[0], ctr[1], ctr[2], ctr[3]}};
  mc_u32x4 r = mc_philox4x32_10(c, key[0], key[1]);
  memcpy(out, r.v, sizeof(r.v));
}
