// Evolution mode drivers of the host drop-in.
#include "evolve.h"

#include <math.h>
#include <stdio.h>
#include <string.h>
#include <strings.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "io.h"

namespace mc {
namespace {

void check(int rc, const char *what) {
  if (rc != MC_OK) throw std::runtime_error(std::string(what) + ": " + mc_last_error());
}

int potential_id(const char *name) {
  if (strcasecmp(name, "ll") == 0 || strcasecmp(name, "lebwohl-lahser") == 0) return MC_LL;
  if (strcasecmp(name, "ghrl") == 0 || strcasecmp(name, "grun-hess") == 0) return MC_GHRL;
  if (strcasecmp(name, "pear") == 0) return MC_PEAR;
  throw ConfigError{2, std::string("Potential '") + name + "' not implemented. Options: ll | ghrl | pear"};
}

// Largest eigenvalue of a symmetric 3x3 matrix {xx,xy,xz,yy,yz,zz} by the
// trigonometric closed form (Smith, CACM 4(4) 1961).
double sym3_max(const double q[6]) {
  const double a = q[0], b = q[1], c = q[2], d = q[3], e = q[4], f = q[5];
  const double p1 = b * b + c * c + e * e;
  const double tr = (a + d + f) / 3.0;
  const double p2 = (a - tr) * (a - tr) + (d - tr) * (d - tr) + (f - tr) * (f - tr) + 2.0 * p1;
  if (p2 <= 1e-300) return tr;
  const double pp = sqrt(p2 / 6.0);
  const double B0 = (a - tr) / pp, B1 = b / pp, B2 = c / pp, B3 = (d - tr) / pp, B4 = e / pp, B5 = (f - tr) / pp;
  const double det = B0 * (B3 * B5 - B4 * B4) - B1 * (B1 * B5 - B4 * B2) + B2 * (B1 * B4 - B3 * B2);
  double r = det / 2.0;
  r = r < -1.0 ? -1.0 : (r > 1.0 ? 1.0 : r);
  const double phi = acos(r) / 3.0;
  const double l0 = tr + 2.0 * pp * cos(phi);
  const double l2 = tr + 2.0 * pp * cos(phi + 2.0 * M_PI / 3.0);
  const double l1 = 3.0 * tr - l0 - l2;
  return fmax(l0, fmax(l1, l2));
}

inline bool resolve(int bound, int &ii, int N) {
  if (bound == MC_PERIODIC) {
    ii = (ii + N) % N;
    return true;
  }
  return !(ii == -1 || ii == N);
}

// Number of rows of the reference's float loops
// for(x=xi; (int)1e6*sign*(x-xf)>=0; x+=dx) with sign = (int)(-dx/|dx|).
template <typename F>
void float_loop(float xi, float xf, float dx, F body) {
  if (dx == 0) throw ConfigError{2, "the loop step (dT or dE) must not be zero"};
  const int sign = (int)(-dx / fabsf(dx));
  for (float x = xi; (float)(1000000 * sign) * (x - xf) >= 0; x += dx) body(x);
}

}  // namespace

void Evolve::check_Points(int *pt, const Parameters &p) {
  VallidPoints = 0;
  const long long nt = (long long)p.Nx * p.Ny * p.Nz;
  for (long long s = 0; s < nt; ++s)
    if (pt[s]) ++VallidPoints;
}

EvolveNGPU::EvolveNGPU(float *ni_, int *pt_, Parameters *params_)
    : ni(ni_), pt(pt_), params(params_), Nx(params_->Nx), Ny(params_->Ny), Nz(params_->Nz) {}

EvolveNGPU::~EvolveNGPU() {
  if (h) mc_destroy(h);
}

void EvolveNGPU::open() {
  if (h) return;
  if (!geometry) throw std::runtime_error("Evolve::run called before a geometry was attached");
  mc_params mp;
  memset(&mp, 0, sizeof(mp));
  mp.nx = Nx;
  mp.ny = Ny;
  mp.nz = Nz;
  for (int a = 0; a < 3; ++a) mp.bound[a] = geometry->bound[a];
  mp.potential = potential_id(params->potential);
  mp.A = params->A;
  mp.B1 = params->B1;
  mp.B2 = params->B2;
  mp.C = params->C;
  mp.ghrl_rho = params->ghrl_rho;
  mp.ghrl_lambda = params->ghrl_lambda;
  mp.ghrl_mu = params->ghrl_mu;
  mp.ghrl_nu = params->ghrl_nu;
  mp.ghrl_sigma = params->ghrl_sigma;
  mp.elecX = params->elecX;
  mp.elecY = params->elecY;
  mp.elecZ = params->elecZ;
  mp.elecA = params->elecA;
  mp.elecE = params->elecE;
  mp.T = params->T;
  if ((int)geometry->surfaces.size() > MC_MAX_SURF) throw ConfigError{2, "at most 8 anchoring surfaces"};
  mp.n_surf = (int)geometry->surfaces.size();
  for (int s = 0; s < mp.n_surf; ++s) {
    mp.surf_kind[s] = geometry->surfaces[s].kind;
    mp.surf_W[s] = geometry->surfaces[s].W;
    for (int c = 0; c < 3; ++c) mp.surf_dir[s][c] = geometry->surfaces[s].n_s[c];
  }
  mp.seed = (uint64_t)(uint32_t)params->seed;
  mp.stream = 0;
  mp.order = cpu_semantics ? MC_ORDER_REF8 : MC_ORDER_CHECKER;
  mp.ang_var = 0.5f;  // every reference run() starts at 0.5
  check(mc_create(&mp, ni, pt, mp.n_surf ? geometry->ns.data() : nullptr, &h), "mc_create");
}

void EvolveNGPU::sweeps(int n) {
  if (n > 0) check(mc_sweeps_graph(h, n), "mc_sweeps_graph");
}

void EvolveNGPU::measured(int n, double &E, double &E2, double &S, double &S2) {
  if (n <= 0) {
    volatile double zero = 0.0;
    E = E2 = S = S2 = zero / zero;  // the reference divides by MCS = 0 too
    return;
  }
  std::vector<double> Es(n), Ss(n);
  check(mc_measure_series(h, n, Es.data(), Ss.data()), "mc_measure_series");
  E = E2 = S = S2 = 0;
  for (int w = 0; w < n; ++w) {
    E += Es[w];
    E2 += Es[w] * Es[w];
    S += Ss[w];
    S2 += Ss[w] * Ss[w];
  }
  E /= n;
  E2 /= n;
  S /= n;
  S2 /= n;
}

void EvolveNGPU::set_T(float T) {
  params->T = T;
  check(mc_set_T(h, T), "mc_set_T");
}

void EvolveNGPU::set_field(float E) {
  params->elecE = E;
  check(mc_set_field(h, E), "mc_set_field");
}

void EvolveNGPU::fetch() { check(mc_get_directors_aos(h, ni), "mc_get_directors_aos"); }

void EvolveNGPU::snapshot(const std::string &fname) {
  fetch();
  // The local order parameter of every site comes from the device (one
  // kernel) instead of Nt host eigen-solves.
  std::vector<float> S((size_t)Nx * Ny * Nz);
  check(mc_local_order(h, S.data()), "mc_local_order");
  print_n(fname.c_str(), ni, *params, pt, geometry->bound, S.data());
}

FILE *EvolveNGPU::open_po(const char *first_col) {
  FILE *f = fopen("po.dat", "a");
  if (!f) throw std::runtime_error("cannot open po.dat");
  if (strcasecmp(params->potential, "pear") == 0)
    fprintf(f, "%s S varS E varE P\n", first_col);
  else
    fprintf(f, "%s S varS E varE\n", first_col);
  fflush(f);
  return f;
}

void EvolveNGPU::row(FILE *f, const char *x, double E, double E2, double S1, double S2) {
  if (strcasecmp(params->potential, "pear") == 0)
    fprintf(f, "%s %g %g %g %g %g\n", x, S1, S2 - S1 * S1, E, E2 - E * E, polarization());
  else
    fprintf(f, "%s %g %g %g %g\n", x, S1, S2 - S1 * S1, E, E2 - E * E);
  fflush(f);
}

// src/parameter_order.cpp:142-155 (over every site, divided by Nt), FP64.
double EvolveNGPU::polarization() const {
  const long long nt = (long long)Nx * Ny * Nz;
  double P[3] = {0, 0, 0};
  for (long long s = 0; s < nt; ++s)
    for (int c = 0; c < 3; ++c) P[c] += ni[3 * s + c];
  for (int c = 0; c < 3; ++c) P[c] /= (double)nt;
  return sqrt(P[0] * P[0] + P[1] * P[1] + P[2] * P[2]);
}

// src/evolve_thermal.cu:33-116 (CPU twin src/evolve_thermal.cpp:30-95).
int thermalEvolveNGPU::run() {
  params->T = params->Ti;
  open();
  FILE *po = open_po("T");
  printf("Starting thermal variation from %g to %g with step %g\nMCT=%d MCS=%d\n", params->Ti, params->Tf,
         params->dT, params->MCT, params->MCS);
  fflush(stdout);
  float_loop(params->Ti, params->Tf, params->dT, [&](float T) {
    set_T(T);
    sweeps(params->MCT);
    double E, E2, S1, S2;
    measured(params->MCS, E, E2, S1, S2);
    char name[64], x[32];
    snprintf(name, sizeof(name), "director_field_%d.csv",
             cpu_semantics ? (int)(100 * (T + 1e-7)) : (int)(100 * (T + 1e-5)));
    snapshot(name);
    snprintf(x, sizeof(x), "%g", (double)T);
    row(po, x, E, E2, S1, S2);
  });
  fclose(po);
  return 0;
}

// src/evolve_step.cu:31-115.
int stepEvolveNGPU::run() {
  params->T = params->Ti;
  open();
  FILE *po = open_po("ii");
  printf("Step relaxation using MCT=%d MCS=%d and fn=%d\n", params->MCT, params->MCS, params->fn);
  fflush(stdout);
  for (int ii = params->first_file; ii < params->fn + params->first_file; ++ii) {
    sweeps(params->MCT);
    double E, E2, S1, S2;
    measured(params->MCS, E, E2, S1, S2);
    snapshot("director_field_" + std::to_string(ii) + ".csv");
    row(po, std::to_string(ii).c_str(), E, E2, S1, S2);
  }
  fclose(po);
  return 0;
}

// src/evolve_quench.cu:31-119: MCT sweeps at Ti, then (int)(MCT*dT) at Tf
// (the CPU build runs every integer below MCT*dT, src/evolve_quench.cpp:62).
int quenchEvolveNGPU::run() {
  params->T = params->Ti;
  open();
  FILE *po = open_po("ii");
  const float prod = params->MCT * params->dT;
  int second = cpu_semantics ? (prod > 0 ? (int)ceilf(prod) : 0) : (int)prod;
  if (second < 0) second = 0;
  printf("Quench: MCT=%d at T=%g, %d sweeps and MCS=%d at T=%g, fn=%d\n", params->MCT, params->Ti, second,
         params->MCS, params->Tf, params->fn);
  fflush(stdout);
  for (int ii = params->first_file; ii < params->fn + params->first_file; ++ii) {
    set_T(params->Ti);
    sweeps(params->MCT);
    set_T(params->Tf);
    sweeps(second);
    double E, E2, S1, S2;
    measured(params->MCS, E, E2, S1, S2);
    snapshot("director_field_" + std::to_string(ii) + ".csv");
    row(po, std::to_string(ii).c_str(), E, E2, S1, S2);
  }
  fclose(po);
  return 0;
}

// src/evolve_electric.cu:33-118.
int electricEvolveNGPU::run() {
  params->T = params->Ti;
  params->elecE = params->elecEi;
  open();
  FILE *po = open_po("T");
  printf("Starting electric variation from %g to %g with step %g\nMCT=%d MCS=%d\n", params->elecEi,
         params->elecEf, params->elecdE, params->MCT, params->MCS);
  fflush(stdout);
  float_loop(params->elecEi, params->elecEf, params->elecdE, [&](float E) {
    set_field(E);
    sweeps(params->MCT);
    double Em, E2, S1, S2;
    measured(params->MCS, Em, E2, S1, S2);
    char name[64], x[32];
    snprintf(name, sizeof(name), "director_field_%d.csv",
             cpu_semantics ? (int)(100 * (E + 1e-7)) : (int)(1000 * (E + 1e-5)));
    snapshot(name);
    snprintf(x, sizeof(x), "%g", (double)E);
    row(po, x, Em, E2, S1, S2);
  });
  fclose(po);
  return 0;
}

double lattice_order_parameter(const float *ni, const int *pt, int i, int j, int k, const Parameters &p,
                               const int bound[3]) {
  const int Nx = p.Nx, Ny = p.Ny, Nz = p.Nz;
  double Q[6] = {0, 0, 0, 0, 0, 0};
  int points = 0;
  static const int off[7][3] = {{-1, 0, 0}, {0, -1, 0}, {0, 0, -1}, {0, 0, 0}, {0, 0, 1}, {0, 1, 0}, {1, 0, 0}};
  for (const auto &d : off) {
    int ii = i + d[0], jj = j + d[1], kk = k + d[2];
    if (!resolve(bound[0], ii, Nx) || !resolve(bound[1], jj, Ny) || !resolve(bound[2], kk, Nz)) continue;
    const long long t = (long long)ii + (long long)Nx * (jj + (long long)Ny * kk);
    if (!pt[t]) continue;
    const double x = ni[3 * t], y = ni[3 * t + 1], z = ni[3 * t + 2];
    Q[0] += 3 * x * x - 1;
    Q[1] += 3 * x * y;
    Q[2] += 3 * x * z;
    Q[3] += 3 * y * y - 1;
    Q[4] += 3 * y * z;
    Q[5] += 3 * z * z - 1;
    ++points;
  }
  for (double &q : Q) q /= 2.0 * points;
  return sym3_max(Q);
}

int print_n(const char *fname, const float *ni, const Parameters &p, const int *pt, const int bound[3],
            const float *S_site) {
  FILE *out = fopen(fname, "w");
  if (!out) {
    perror(fname);
    return 1;
  }
  fputs("x,y,z,nx,ny,nz,S,pt\n", out);
  const long long nt = (long long)p.Nx * p.Ny * p.Nz;
  const long long chunk = 1 << 16;
  const long long nchunks = (nt + chunk - 1) / chunk;
  // Format chunks in parallel, write them in order.
  const int batch = 64;
  std::vector<std::string> bufs(batch);
  for (long long c0 = 0; c0 < nchunks; c0 += batch) {
    const long long c1 = std::min(nchunks, c0 + batch);
#pragma omp parallel for schedule(dynamic, 1)
    for (long long c = c0; c < c1; ++c) {
      std::string &b = bufs[c - c0];
      b.clear();
      char line[160];
      const long long a = c * chunk, e = std::min(nt, a + chunk);
      for (long long s = a; s < e; ++s) {
        const int i = (int)(s % p.Nx), j = (int)((s / p.Nx) % p.Ny), k = (int)(s / ((long long)p.Nx * p.Ny));
        const float S = !pt[s] ? (float)NAN
                        : S_site ? S_site[s] : (float)lattice_order_parameter(ni, pt, i, j, k, p, bound);
        const int n = snprintf(line, sizeof(line), "%d,%d,%d,%g,%g,%g,%g,%d\n", i, j, k, (double)ni[3 * s],
                               (double)ni[3 * s + 1], (double)ni[3 * s + 2], (double)S, pt[s]);
        b.append(line, n);
      }
    }
    for (long long c = c0; c < c1; ++c) fwrite(bufs[c - c0].data(), 1, bufs[c - c0].size(), out);
  }
  printf("  Snapshot -> %s\n", fname);
  fflush(stdout);
  fclose(out);
  return 0;
}

}  // namespace mc
