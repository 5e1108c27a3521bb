/* Exercises the CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer:
 * sweeps in both orders for every potential, bulk and slab geometries with each
 * anchoring kind, injected passes, energies, Q tensor and eigenvalues, the local
 * order parameter at every site (faces and corners included), the float loop,
 * expf, Philox and the MLP oracle's forward pass and gradient.  Built by
 * tests/test_sanitize_cpu.py with gcc -fsanitize=address,undefined; exits 0
 * when every call returned without a sanitizer report. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../oracle/mc_oracle.h"
#include "../../oracle/mlp_oracle.h"

static void fill_unit(float *n, int nt, unsigned seed) {
  srand(seed);
  for (int s = 0; s < nt; ++s) {
    double a = rand() / (double)RAND_MAX - 0.5, b = rand() / (double)RAND_MAX - 0.5, c = rand() / (double)RAND_MAX - 0.5;
    double r = sqrt(a * a + b * b + c * c) + 1e-12;
    n[3 * s] = (float)(a / r);
    n[3 * s + 1] = (float)(b / r);
    n[3 * s + 2] = (float)(c / r);
  }
}

static int run_lattice(int nx, int ny, int nz, int pot, int zfree, int kind) {
  const int nt = nx * ny * nz;
  orc_params p;
  memset(&p, 0, sizeof p);
  p.nx = nx;
  p.ny = ny;
  p.nz = nz;
  p.bound[0] = p.bound[1] = ORC_PERIODIC;
  p.bound[2] = zfree ? ORC_FREE : ORC_PERIODIC;
  p.potential = pot;
  p.A = 1.0f;
  p.B1 = 0.3f;
  p.C = 0.3f;
  if (pot == ORC_GHRL) orc_set_ghrl(&p, 1.0f, 0.7f, 1.3f, 16.0f);
  p.elecZ = 1.0f;
  p.elecA = 0.8f;
  p.elecE = 0.5f;
  float *n = malloc(sizeof(float) * 3 * nt), *ns = calloc(3 * (size_t)nt, sizeof(float));
  int32_t *pt = malloc(sizeof(int32_t) * nt);
  fill_unit(n, nt, (unsigned)(nx * 131 + pot * 7 + kind));
  for (int s = 0; s < nt; ++s) pt[s] = 1;
  if (zfree && kind >= 0) {
    p.n_surf = 2;
    for (int f = 0; f < 2; ++f) {
      p.surf_kind[f] = kind;
      p.surf_W[f] = 2.0f;
      p.surf_dir[f][0] = 0.6f;
      p.surf_dir[f][2] = 0.8f;
    }
    for (int s = 0; s < nx * ny; ++s) {
      pt[s] = 2;
      ns[3 * s + 2] = -1.0f;
      pt[nt - 1 - s] = 3;
      ns[3 * (nt - 1 - s) + 2] = 1.0f;
    }
  }
  float av = 0.5f;
  for (int w = 0; w < 3; ++w) {
    orc_sweep_rng(&p, n, pt, ns, 0.9f, &av, 7u, 0u, (uint32_t)w, 0u, 0, 2);
    orc_sweep_rng(&p, n, pt, ns, 0.9f, &av, 7u, 0u, (uint32_t)(w + 3), 0u, 1, 1);
  }
  float *prop = malloc(sizeof(float) * 3 * nt), *u = malloc(sizeof(float) * nt);
  float *eo = malloc(sizeof(float) * nt), *en = malloc(sizeof(float) * nt);
  uint8_t *acc = malloc((size_t)nt);
  fill_unit(prop, nt, 99u);
  for (int s = 0; s < nt; ++s) u[s] = (float)((s % 97) / 97.0);
  for (int set = 0; set < 10; ++set) orc_pass_inject(&p, n, pt, ns, set, 0.7f, prop, u, acc, eo, en);
  float *se = malloc(sizeof(float) * nt);
  orc_site_energies(&p, n, pt, ns, se);
  double e = orc_energy_mean(&p, n, pt, ns, 1, 2) + orc_energy_mean(&p, n, pt, ns, 0, 1);
  double q6[6], ev[3];
  orc_qtensor(n, pt, nt, q6);
  orc_sym3_eigvals(q6, ev);
  double ls = 0;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) ls += orc_local_S(&p, n, pt, i, j, k);
  int ok = isfinite(e) && isfinite(ev[2]) && isfinite(ls);
  free(n);
  free(ns);
  free(pt);
  free(prop);
  free(u);
  free(eo);
  free(en);
  free(acc);
  free(se);
  return ok;
}

int main(void) {
  int ok = 1;
  for (int pot = 0; pot < 3; ++pot) {
    ok &= run_lattice(6, 4, 8, pot, 0, -1);
    ok &= run_lattice(5, 7, 4, pot, 1, -1);
    for (int kind = 0; kind < 8; ++kind) ok &= run_lattice(4, 6, 5, pot, 1, kind);
  }
  float xs[64];
  ok &= orc_float_loop(1.2f, 0.0f, -0.02f, xs, 64) == 61;
  float x[9] = {-100.0f, -88.0f, -10.5f, -1e-3f, 0.0f, 1e-3f, 10.0f, 88.5f, INFINITY}, y[9];
  orc_expf_n(x, y, 9);
  uint32_t ctr[4] = {1, 2, 3, 4}, key[2] = {5, 6}, out[4];
  orc_philox4x32_10(ctr, key, out);
  float *params = malloc(sizeof(float) * 5892);
  for (int i = 0; i < 5892; ++i) params[i] = 0.01f * (float)((i * 37) % 101 - 50) / 50.0f;
  float nbr[18], field[3] = {0.0f, 0.0f, 0.5f}, feat[22], base[3];
  fill_unit(nbr, 6, 3u);
  orc_mlp_features(nbr, field, 1.1f, feat, base);
  double o4[4], mu[3], lk;
  orc_mlp_forward(params, feat, 1, o4);
  orc_mlp_head(o4, base, mu, &lk);
  float X[4 * 22], B[4 * 3], D[4 * 3];
  for (int b = 0; b < 4; ++b) {
    memcpy(X + 22 * b, feat, sizeof feat);
    memcpy(B + 3 * b, base, sizeof base);
    memcpy(D + 3 * b, nbr + 3 * b, 3 * sizeof(float));
  }
  double *g = calloc(5892, sizeof(double));
  double loss = orc_mlp_nll_grad(params, 4, X, B, D, 1, g);
  ok &= isfinite(loss) && isfinite(lk);
  free(params);
  free(g);
  printf("oracle sanitizer run %s\n", ok ? "ok" : "produced non-finite results");
  return ok ? 0 : 1;
}
