// Evolution mode drivers of the host drop-in.
#include "evolve.h"

#include <ctype.h>
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
// The body returns true to end the loop early.
template <typename F>
void float_loop(float xi, float xf, float dx, F body) {
  if (dx == 0) throw ConfigError{2, "the loop step (dT or dE) must not be zero"};
  const int sign = (int)(-dx / fabsf(dx));
  for (float x = xi; (float)(1000000 * sign) * (x - xf) >= 0; x += dx)
    if (body(x)) break;
}

// 64-bit FNV-1a.
struct Fnv {
  uint64_t h = 0xcbf29ce484222325ull;
  void add(const void *p, size_t n) {
    const unsigned char *b = static_cast<const unsigned char *>(p);
    for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 0x100000001b3ull;
  }
  void add_str(const char *s) { add(s, strlen(s) + 1); }
  template <typename T>
  void add_val(T v) {
    add(&v, sizeof(v));
  }
};

// Checkpoint file, little-endian:
//   "MCLCKPT1", u32 version (1), i32 Nx, Ny, Nz, char evol[20],
//   u64 fingerprint of the run's parameters and lattice, u32 index of the
//   next sweep, f32 ang_var, i32 rows done, f32 loop variable of the next
//   row, f32 directors[3*Nt] (AoS), then u64 FNV-1a of all bytes before it.
constexpr char kMagic[8] = {'M', 'C', 'L', 'C', 'K', 'P', 'T', '1'};
constexpr uint32_t kCkptVersion = 1;
constexpr size_t kCkptHead = 8 + 4 + 12 + 20 + 8 + 4 + 4 + 4 + 4;

template <typename T>
void put(std::vector<unsigned char> &b, T v) {
  const unsigned char *p = reinterpret_cast<const unsigned char *>(&v);
  b.insert(b.end(), p, p + sizeof(T));
}
template <typename T>
T get(const unsigned char *&p) {
  T v;
  memcpy(&v, p, sizeof(T));
  p += sizeof(T);
  return v;
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
  if (rows_done > 0) return f;  // a resumed run appends to the rows already there
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

// What the chain and the rows written so far depend on: mode, lattice,
// boundaries, potential and its constants, field, anchoring, point types and
// surface normals, seed, sweep counts and loop start and step.  The loop ends
// (Tf, elecEf, fn) are left out, so a resumed run may go further or stop
// earlier than the one that wrote the file.
uint64_t EvolveNGPU::chain_fingerprint() const {
  Fnv f;
  const Parameters &p = *params;
  char evol[20] = {0};
  snprintf(evol, sizeof(evol), "%s", p.evol);
  for (char &c : evol) c = (char)tolower((unsigned char)c);
  f.add_str(evol);
  f.add_str(geometry->name());
  f.add_str(p.potential);
  for (int v : {p.Nx, p.Ny, p.Nz, geometry->bound[0], geometry->bound[1], geometry->bound[2], p.MCT, p.MCS, p.seed,
                p.first_file, cpu_semantics ? 1 : 0})
    f.add_val(v);
  for (float v : {p.A, p.B1, p.B2, p.C, p.ghrl_rho, p.ghrl_lambda, p.ghrl_mu, p.ghrl_nu, p.ghrl_sigma, p.elecX,
                  p.elecY, p.elecZ, p.elecA, p.elecEi, p.elecdE, p.Ti, p.dT})
    f.add_val(v);
  for (const Anchoring &a : geometry->surfaces) {
    f.add_val(a.kind);
    f.add_val(a.W);
    f.add(a.n_s, sizeof(a.n_s));
  }
  const size_t nt = (size_t)Nx * Ny * Nz;
  f.add(pt, nt * sizeof(int));
  if (!geometry->ns.empty()) f.add(geometry->ns.data(), geometry->ns.size() * sizeof(float));
  return f.h;
}

void EvolveNGPU::start_run(float *x) {
  rows_done = rows_this_run = 0;
  const size_t nt = (size_t)Nx * Ny * Nz;
  std::vector<unsigned char> b;
  if (!control.resume.empty()) {
    // Read and check the whole file before the device is touched.
    FILE *f = fopen(control.resume.c_str(), "rb");
    if (!f) throw ConfigError{2, "cannot open checkpoint " + control.resume};
    unsigned char chunk[1 << 16];
    size_t got;
    while ((got = fread(chunk, 1, sizeof(chunk), f)) > 0) b.insert(b.end(), chunk, chunk + got);
    fclose(f);
    const std::string what = "checkpoint " + control.resume + ": ";
    if (b.size() != kCkptHead + 12 * nt + 8)
      throw ConfigError{2, what + "its size does not fit a " + std::to_string(Nx) + "x" + std::to_string(Ny) + "x" +
                               std::to_string(Nz) + " lattice"};
    Fnv sum;
    sum.add(b.data(), b.size() - 8);
    uint64_t stored;
    memcpy(&stored, b.data() + b.size() - 8, 8);
    if (memcmp(b.data(), kMagic, 8) != 0) throw ConfigError{2, what + "not a checkpoint file"};
    if (stored != sum.h) throw ConfigError{2, what + "checksum mismatch (file damaged)"};
    const unsigned char *q = b.data() + 8;
    if (get<uint32_t>(q) != kCkptVersion) throw ConfigError{2, what + "unknown version"};
    const int nx = get<int32_t>(q), ny = get<int32_t>(q), nz = get<int32_t>(q);
    q += 20;  // evol: covered by the fingerprint
    const uint64_t fp = get<uint64_t>(q);
    if (nx != Nx || ny != Ny || nz != Nz || fp != chain_fingerprint())
      throw ConfigError{2, what + "written by a run with other parameters, lattice or mode"};
  }
  open();
  if (b.empty()) return;
  const unsigned char *q = b.data() + 8 + 4 + 12 + 20 + 8;
  const uint32_t sweep = get<uint32_t>(q);
  const float av = get<float>(q);
  const int rows = get<int32_t>(q);
  const float nx = get<float>(q);
  memcpy(ni, q, 12 * nt);
  check(mc_set_directors_aos(h, ni), "mc_set_directors_aos");
  check(mc_set_ang_var(h, av), "mc_set_ang_var");
  check(mc_set_sweep_index(h, sweep), "mc_set_sweep_index");
  rows_done = rows;
  *x = nx;
  printf("Resuming from %s: %d row(s) done, next sweep %u, ang_var %g\n", control.resume.c_str(), rows, sweep,
         (double)av);
  fflush(stdout);
}

void EvolveNGPU::write_checkpoint(float next_x) {
  const size_t nt = (size_t)Nx * Ny * Nz;
  fetch();
  float av = 0;
  uint32_t acc = 0, sweep = 0;
  check(mc_get_ang_var(h, &av), "mc_get_ang_var");
  check(mc_get_acceptance(h, &acc, &sweep), "mc_get_acceptance");
  std::vector<unsigned char> b;
  b.reserve(kCkptHead + 12 * nt + 8);
  b.insert(b.end(), kMagic, kMagic + 8);
  put<uint32_t>(b, kCkptVersion);
  put<int32_t>(b, Nx);
  put<int32_t>(b, Ny);
  put<int32_t>(b, Nz);
  char evol[20] = {0};
  snprintf(evol, sizeof(evol), "%s", params->evol);
  b.insert(b.end(), evol, evol + 20);
  put<uint64_t>(b, chain_fingerprint());
  put<uint32_t>(b, sweep);
  put<float>(b, av);
  put<int32_t>(b, rows_done);
  put<float>(b, next_x);
  const unsigned char *d = reinterpret_cast<const unsigned char *>(ni);
  b.insert(b.end(), d, d + 12 * nt);
  Fnv sum;
  sum.add(b.data(), b.size());
  put<uint64_t>(b, sum.h);
  // Written beside the target and renamed over it, so an interrupted write
  // leaves the previous checkpoint whole.
  const std::string tmp = control.checkpoint + ".tmp";
  FILE *f = fopen(tmp.c_str(), "wb");
  if (!f) throw std::runtime_error("cannot write checkpoint " + tmp);
  const bool ok = fwrite(b.data(), 1, b.size(), f) == b.size();
  if (fclose(f) != 0 || !ok) throw std::runtime_error("short write of checkpoint " + tmp);
  if (rename(tmp.c_str(), control.checkpoint.c_str()) != 0)
    throw std::runtime_error("cannot rename " + tmp + " to " + control.checkpoint);
}

bool EvolveNGPU::after_row(float next_x) {
  ++rows_done;
  ++rows_this_run;
  if (!control.checkpoint.empty()) write_checkpoint(next_x);
  if (control.max_rows > 0 && rows_this_run >= control.max_rows) {
    printf("Stopping after %d row(s) of this run (%d in all)%s%s\n", rows_this_run, rows_done,
           control.checkpoint.empty() ? "" : "; checkpoint ", control.checkpoint.c_str());
    fflush(stdout);
    return true;
  }
  return false;
}

// src/evolve_thermal.cu:33-116 (CPU twin src/evolve_thermal.cpp:30-95).
int thermalEvolveNGPU::run() {
  params->T = params->Ti;
  float T0 = params->Ti;
  start_run(&T0);
  FILE *po = open_po("T");
  printf("Starting thermal variation from %g to %g with step %g\nMCT=%d MCS=%d\n", params->Ti, params->Tf,
         params->dT, params->MCT, params->MCS);
  fflush(stdout);
  float_loop(T0, params->Tf, params->dT, [&](float T) {
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
    return after_row(T + params->dT);
  });
  fclose(po);
  return 0;
}

// src/evolve_step.cu:31-115.
int stepEvolveNGPU::run() {
  params->T = params->Ti;
  float unused = 0;
  start_run(&unused);
  FILE *po = open_po("ii");
  printf("Step relaxation using MCT=%d MCS=%d and fn=%d\n", params->MCT, params->MCS, params->fn);
  fflush(stdout);
  for (int ii = params->first_file + rows_done; ii < params->fn + params->first_file; ++ii) {
    sweeps(params->MCT);
    double E, E2, S1, S2;
    measured(params->MCS, E, E2, S1, S2);
    snapshot("director_field_" + std::to_string(ii) + ".csv");
    row(po, std::to_string(ii).c_str(), E, E2, S1, S2);
    if (after_row(0.0f)) break;
  }
  fclose(po);
  return 0;
}

// src/evolve_quench.cu:31-119: MCT sweeps at Ti, then (int)(MCT*dT) at Tf
// (the CPU build runs every integer below MCT*dT, src/evolve_quench.cpp:62).
int quenchEvolveNGPU::run() {
  params->T = params->Ti;
  float unused = 0;
  start_run(&unused);
  FILE *po = open_po("ii");
  const float prod = params->MCT * params->dT;
  int second = cpu_semantics ? (prod > 0 ? (int)ceilf(prod) : 0) : (int)prod;
  if (second < 0) second = 0;
  printf("Quench: MCT=%d at T=%g, %d sweeps and MCS=%d at T=%g, fn=%d\n", params->MCT, params->Ti, second,
         params->MCS, params->Tf, params->fn);
  fflush(stdout);
  for (int ii = params->first_file + rows_done; ii < params->fn + params->first_file; ++ii) {
    set_T(params->Ti);
    sweeps(params->MCT);
    set_T(params->Tf);
    sweeps(second);
    double E, E2, S1, S2;
    measured(params->MCS, E, E2, S1, S2);
    snapshot("director_field_" + std::to_string(ii) + ".csv");
    row(po, std::to_string(ii).c_str(), E, E2, S1, S2);
    if (after_row(0.0f)) break;
  }
  fclose(po);
  return 0;
}

// src/evolve_electric.cu:33-118.
int electricEvolveNGPU::run() {
  params->T = params->Ti;
  params->elecE = params->elecEi;
  float E0 = params->elecEi;
  start_run(&E0);
  FILE *po = open_po("T");
  printf("Starting electric variation from %g to %g with step %g\nMCT=%d MCS=%d\n", params->elecEi,
         params->elecEf, params->elecdE, params->MCT, params->MCS);
  fflush(stdout);
  float_loop(E0, params->elecEf, params->elecdE, [&](float E) {
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
    return after_row(E + params->elecdE);
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
