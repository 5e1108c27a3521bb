// C ABI of the checkerboard Metropolis engine (declared in include/mclics_hip.h).
//
// One handle owns one lattice on the device that was current at mc_create, a
// private HIP stream (default priority) and the small device state that makes
// sweeps host-sync free.  Errors are returned as codes with a thread-local
// message; no function terminates the process.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <string>
#include <vector>

#include "../../include/mclics_hip.h"
#include "mc_kernels.h"
#include "mc_mlp.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define MC_HIP(call)                                                                                       \
  do {                                                                                                     \
    hipError_t e_ = (call);                                                                                \
    if (e_ != hipSuccess) {                                                                                \
      (void)hipGetLastError();                                                                             \
      return fail(e_ == hipErrorOutOfMemory ? MC_ENOMEM : MC_EHIP,                                         \
                  std::string(#call) + " failed: " + hipGetErrorString(e_));                              \
    }                                                                                                      \
  } while (0)

#define MC_LAUNCHED(what)                                                                                  \
  do {                                                                                                     \
    hipError_t e_ = hipGetLastError();                                                                     \
    if (e_ != hipSuccess) return fail(MC_EHIP, std::string(what) + " launch failed: " + hipGetErrorString(e_)); \
  } while (0)

__global__ void k_set_state_f(MCDevState *st, int field, float v) {
  if (field == 0) st->T = v;
  if (field == 1) st->elecE = v;
  if (field == 2) st->ang_var = v;
}

__global__ void k_set_sweep_index(MCDevState *st, uint32_t sweep) { st->sweep = sweep; }

__global__ void k_init_state(MCDevState *st, float T, float elecE, float ang_var, unsigned long long *ctr) {
  for (int c = threadIdx.x; c < MCD_NSUB; c += blockDim.x) ctr[(size_t)c * MCD_SUB_STRIDE] = 0ull;
  if (threadIdx.x != 0) return;
  st->T = T;
  st->elecE = elecE;
  st->ang_var = ang_var;
  st->sweep = 0u;
  st->acc_count = 0u;
  st->blocks_done = 0u;
  st->last_acc = 0u;
  st->meas_done = 0u;
}

}  // namespace

void mc_set_error_message(const char *msg) { g_err = msg; }

int mc_handle_launch(mc_handle *h, MCLaunch *out);

struct mc_handle {
  int device = 0;
  hipStream_t stream = nullptr;
  MCDevParams P{};
  int order = MC_ORDER_CHECKER;
  int kernel = MC_KERNEL_AUTO;  // sweep kernel selection as requested
  int active = MC_KERNEL_GENERIC;  // kernel in use (AUTO resolved)
  int x4_grid = 0;              // workgroups of a segment pass (0: 8 per CU)
  int x4_xcd = 0;               // tile schedule of the segment pass
  mc_mlp *learned = nullptr;    // learned proposal network (null: plain sweeps)
  int learned_every = 0;        // every k-th sweep uses the learned proposal
  int learned_reflect = 1;      // learned sweeps add the reflection move
  float *lrbuf = nullptr;       // learned-sweep record of the reflection move (5 planes)
  uint8_t *lrabuf = nullptr;    // and its decisions
  long long host_sweeps = 0;    // sweeps enqueued so far (selects learned sweeps)
  int cu_count = 256;
  float *nx = nullptr, *ny = nullptr, *nz = nullptr;  // current lattice
  // Second plane set of the slab kernel, which reads one set and writes the
  // other; flip tells which allocation holds the current lattice.
  float *bx = nullptr, *by = nullptr, *bz = nullptr;
  int flip = 0;
  MCSlabShape slab{0, 0, 1};  // ty = 0: default shape
  size_t lds_limit = 65536;
  uint8_t *pt = nullptr;
  float *sx = nullptr, *sy = nullptr, *sz = nullptr;
  MCDevState *st = nullptr;
  unsigned long long *ctr = nullptr;  // sweep sub-counters (P.ctr)
  double *partials = nullptr;
  double *result = nullptr;
  double *series = nullptr;  // per-step {E, S} of mc_measure_series
  int series_cap = 0;
  float *staging = nullptr;  // AoS float[3*Nt]
  int measure_grid = 0;
  bool measure_slab = false;  // measurement on the slab layout (shape mshape)
  MCSlabShape mshape{0, 0, 0};
  // Parity and record buffers, allocated on first use.
  float *pbuf = nullptr;     // u / eo / en / records: 7 float planes
  uint8_t *abuf = nullptr;   // accept flags
  std::map<int, hipGraphExec_t> graphs;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;

  MCLaunch launch() const {
    MCLaunch L;
    L.P = P;
    L.stream = stream;
    L.nx = nx;
    L.ny = ny;
    L.nz = nz;
    L.pt = pt;
    L.sx = sx;
    L.sy = sy;
    L.sz = sz;
    L.st = st;
    return L;
  }
  int grid_for(int total) const {
    long long g = ((long long)total + 255) / 256;
    long long cap = (long long)cu_count * 8;
    if (g > cap) g = cap;
    if (g < 1) g = 1;
    return (int)g;
  }
};

namespace {

void free_handle(mc_handle *h) {
  if (!h) return;
  for (auto &kv : h->graphs) (void)hipGraphExecDestroy(kv.second);
  h->graphs.clear();
  void *ptrs[] = {h->nx, h->ny, h->nz, h->bx, h->by, h->bz, h->pt, h->sx, h->sy, h->sz, h->st, h->ctr, h->partials,
                  h->result, h->staging, h->pbuf, h->abuf, h->series, h->lrbuf, h->lrabuf};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (h->ev0) (void)hipEventDestroy(h->ev0);
  if (h->ev1) (void)hipEventDestroy(h->ev1);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

int check_handle(mc_handle *h) {
  if (!h) return fail(MC_EINVAL, "null handle");
  if (hipSetDevice(h->device) != hipSuccess) return fail(MC_EHIP, "hipSetDevice failed for the handle's device");
  return MC_OK;
}

bool fast_eligible(const mc_handle *h) { return h->order == MC_ORDER_CHECKER && mc_x4_eligible(h->launch()); }

// Default slab tile: the preferred row count that fits, then planes per tile
// halved until there are at least as many tiles as CUs (cz >= 4).
bool slab_default(const mc_handle *h, MCSlabShape *out) {
  const MCDevParams &P = h->P;
  const int prefs[] = {8, 4, 16, 6, 2, 12, 10, 14, 1};
  for (int ty : prefs) {
    MCSlabShape s{ty, P.nz, 1};
    if (!mc_slab_shape_ok(P, s, h->lds_limit)) continue;
    while (s.cz % 2 == 0 && s.cz / 2 >= 4 && (long long)(P.ny / ty) * (P.nz / s.cz) < h->cu_count) s.cz /= 2;
    *out = s;
    return true;
  }
  return false;
}

MCSlabShape slab_shape(const mc_handle *h) {
  if (h->slab.ty > 0) return h->slab;
  MCSlabShape s{1, 1, 1};
  slab_default(h, &s);
  s.xcd = h->slab.xcd;
  return s;
}

bool use_x4(const mc_handle *h) { return h->active == MC_KERNEL_X4; }
bool use_slab(const mc_handle *h) { return h->active == MC_KERNEL_SLAB; }

bool learned_sweep(const mc_handle *h) {
  return h->learned && h->learned_every > 0 && (h->host_sweeps % h->learned_every) == 0;
}

// Sweeps can be captured in graphs: plain sweeps, or learned sweeps only
// (every = 1).  With plain and learned sweeps interleaved, which sweep of a
// graph is learned would depend on where the graph starts.
bool graphable(const mc_handle *h) { return !h->learned || h->learned_every == 1; }

int passes_per_sweep(const mc_handle *h) {
  if (h->order == MC_ORDER_REF8) return 8;
  return use_slab(h) && !learned_sweep(h) ? 1 : 2;
}

void swap_planes(mc_handle *h) {
  std::swap(h->nx, h->bx);
  std::swap(h->ny, h->by);
  std::swap(h->nz, h->bz);
  h->flip ^= 1;
}

void drop_graphs(mc_handle *h) {
  for (auto &kv : h->graphs) (void)hipGraphExecDestroy(kv.second);
  h->graphs.clear();
}

// Resolves the requested kernel and allocates what it needs.
int select_kernel(mc_handle *h, int kind) {
  int want = MC_KERNEL_GENERIC;
  MCSlabShape s;
  const bool slab_ok = fast_eligible(h) && (h->slab.ty > 0 || slab_default(h, &s));
  // AUTO: the slab kernel for Lebwohl-Lasher, whose screened update leaves
  // the sweep bound by bytes and issue; GHRL and pear evaluate every move
  // with the exact FP32/FP64 arithmetic, and the segment kernel's higher
  // occupancy and lack of halo work serve that better (mc_autotune times both).
  if (kind == MC_KERNEL_SLAB || (kind == MC_KERNEL_AUTO && slab_ok && h->P.potential == MCD_LL))
    want = MC_KERNEL_SLAB;
  else if (kind == MC_KERNEL_X4 || (kind == MC_KERNEL_AUTO && fast_eligible(h))) want = MC_KERNEL_X4;
  if (want == MC_KERNEL_SLAB) {
    if (!slab_ok) return fail(MC_EINVAL, "the slab kernel needs a checkerboard order and an all-active lattice, "
                                         "periodic on every axis, with Nx a multiple of 4 and a tile that fits");
    (void)mc_slab_prepare(h->lds_limit);
  }
  if (want == MC_KERNEL_SLAB && !h->bx) {
    const size_t fbytes = (size_t)h->P.nt * sizeof(float);
    float *b[3] = {nullptr, nullptr, nullptr};
    for (int d = 0; d < 3; ++d) {
      if (hipMalloc(&b[d], fbytes) != hipSuccess) {
        (void)hipGetLastError();
        for (int e = 0; e < 3; ++e)
          if (b[e]) (void)hipFree(b[e]);
        if (kind == MC_KERNEL_SLAB) return fail(MC_ENOMEM, "slab kernel: second plane set allocation failed");
        want = MC_KERNEL_X4;
        break;
      }
    }
    if (want == MC_KERNEL_SLAB) {
      h->bx = b[0];
      h->by = b[1];
      h->bz = b[2];
    }
  }
  h->kernel = kind;
  h->active = want;
  return MC_OK;
}

// Pass t (0-based) of one sweep; the last pass adapts ang_var.
void launch_pass(mc_handle *h, int t, const MCRecord *rec, const MCLearnRecord *lrec = nullptr) {
  const MCLaunch L = h->launch();
  const int per = passes_per_sweep(h);
  const int finalize = t == per - 1;
  if (learned_sweep(h)) {
    // one grid for both colours (the sweep counters need it)
    const int slots = std::max(mcd_set_slots(h->P, 8), mcd_set_slots(h->P, 9));
    const int grid = mc_mlp_learned_grid(h->P, slots, h->cu_count);
    mc_mlp_launch_learned_pass(L, mc_mlp_buffers(h->learned), t, grid, finalize, h->learned_reflect, lrec);
  } else if (use_slab(h)) {
    mc_launch_sweep_slab(L, h->bx, h->by, h->bz, slab_shape(h), rec);
    swap_planes(h);
  } else if (use_x4(h)) {
    const int g = h->x4_grid > 0 ? h->x4_grid : h->cu_count * 8;
    mc_launch_sweep_x4(L, t, g, finalize, h->x4_xcd, rec);
  } else {
    const int set = per == 8 ? t : 8 + t;
    int slots = 0;  // one grid for every pass of the sweep (the sweep counters need it)
    for (int u = 0; u < per; ++u) slots = std::max(slots, mc_set_slots(h->P, per == 8 ? u : 8 + u));
    mc_launch_sweep_pass(L, set, h->grid_for(slots), finalize, rec);
  }
}

int enqueue_sweep(mc_handle *h, const MCRecord *rec, const MCLearnRecord *lrec = nullptr) {
  const int per = passes_per_sweep(h);
  for (int t = 0; t < per; ++t) launch_pass(h, t, rec, lrec);
  ++h->host_sweeps;
  MC_LAUNCHED("sweep pass");
  return MC_OK;
}

// E, S and Q of the current lattice into h->result (and series when given):
// one segment-kernel launch for a bulk periodic lattice, else the general
// per-site kernel and its one-block finisher.
void launch_measure(mc_handle *h, double *series) {
  const MCLaunch L = h->launch();
  if (h->measure_slab)  // any sweep order: the measurement reads the lattice only
    mc_launch_measure_slab(L, h->mshape, h->partials, h->result, series);
  else if (mc_x4_eligible(L))
    mc_launch_measure_x4(L, h->measure_grid, h->measure_grid % 8 == 0 ? 1 : 0, h->partials, h->result, series);
  else
    mc_launch_measure(L, h->measure_grid, h->partials, h->result, series);
}

// Graph of n sweeps starting from the current plane set.  Capture
// records the launches without running them, so the host-side plane swaps
// and sweep count made while capturing are undone afterwards.
int sweep_graph(mc_handle *h, int n, hipGraphExec_t *out) {
  const int key = 4 * n + 2 * (h->learned ? 1 : 0) + h->flip;
  auto it = h->graphs.find(key);
  if (it != h->graphs.end()) {
    *out = it->second;
    return MC_OK;
  }
  const int flip0 = h->flip;
  const long long sweeps0 = h->host_sweeps;
  hipGraph_t g = nullptr;
  MC_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  int rc = MC_OK;
  for (int w = 0; w < n && rc == MC_OK; ++w) rc = enqueue_sweep(h, nullptr);
  if (h->flip != flip0) swap_planes(h);
  h->host_sweeps = sweeps0;
  hipError_t e = hipStreamEndCapture(h->stream, &g);
  if (rc != MC_OK || e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    if (rc != MC_OK) return rc;
    MC_HIP(e);
  }
  hipGraphExec_t ge = nullptr;
  e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  MC_HIP(e);
  h->graphs.emplace(key, ge);
  *out = ge;
  return MC_OK;
}

// Replays the graph of n sweeps and advances the host-side bookkeeping the
// way the captured launches advance the device.
int launch_sweep_graph(mc_handle *h, int n, hipGraphExec_t ge) {
  MC_HIP(hipGraphLaunch(ge, h->stream));
  if (use_slab(h) && !h->learned && (n & 1)) swap_planes(h);  // learned sweeps keep the plane set
  h->host_sweeps += n;
  return MC_OK;
}

// n sweeps as graphs of at most kGraphChunk sweeps: the graph of kGraphChunk
// sweeps replayed n / kGraphChunk times, then the graph of the n % kGraphChunk
// that remain.  No graph holds more than kGraphChunk sweeps (8 * kGraphChunk
// kernel nodes with the site kernel's eight passes), and a handle caches at
// most 2 * kGraphChunk graphs (one per length and plane set).  kGraphChunk
// is even, so every replay of the full chunk starts from the plane set the
// remainder graph was captured on.
constexpr int kGraphChunk = 64;
static_assert(kGraphChunk % 2 == 0, "the chunk must keep the slab kernel's plane set");

struct GraphPlan {
  hipGraphExec_t chunk = nullptr, rest = nullptr;
  int reps = 0, rem = 0;
};

int plan_sweep_graphs(mc_handle *h, int n, GraphPlan *g) {
  g->reps = n / kGraphChunk;
  g->rem = n % kGraphChunk;
  int rc = MC_OK;
  if (g->reps) rc = sweep_graph(h, kGraphChunk, &g->chunk);
  if (rc == MC_OK && g->rem) rc = sweep_graph(h, g->rem, &g->rest);
  return rc;
}

int launch_sweep_graphs(mc_handle *h, const GraphPlan &g) {
  for (int r = 0; r < g.reps; ++r) {
    const int rc = launch_sweep_graph(h, kGraphChunk, g.chunk);
    if (rc) return rc;
  }
  return g.rem ? launch_sweep_graph(h, g.rem, g.rest) : MC_OK;
}

int ensure_parity_buffers(mc_handle *h) {
  const size_t nt = (size_t)h->P.nt;
  if (!h->pbuf) MC_HIP(hipMalloc(&h->pbuf, 7 * nt * sizeof(float)));
  if (!h->abuf) MC_HIP(hipMalloc(&h->abuf, nt));
  return MC_OK;
}

}  // namespace

const char *mc_to_dev_params(const mc_params *p, MCDevParams *out) {
  MCDevParams &P = *out;
  P = MCDevParams{};
  P.nx = p->nx;
  P.ny = p->ny;
  P.nz = p->nz;
  for (int a = 0; a < 3; ++a) P.per[a] = p->bound[a] == MC_PERIODIC;
  P.potential = p->potential;
  P.A = p->A;
  P.B1 = p->B1;
  P.C = p->C;
  P.rho = p->ghrl_rho;
  P.lambda = p->ghrl_lambda;
  P.mu = p->ghrl_mu;
  P.nu = p->ghrl_nu;
  P.sigma = p->ghrl_sigma;
  P.eX = p->elecX;
  P.eY = p->elecY;
  P.eZ = p->elecZ;
  P.elecA = p->elecA;
  P.n_surf = p->n_surf;
  for (int s = 0; s < MC_MAX_SURF; ++s) {
    P.surf_kind[s] = s < p->n_surf ? p->surf_kind[s] : 0;
    P.surf_W[s] = s < p->n_surf ? p->surf_W[s] : 0.0f;
    for (int c = 0; c < 3; ++c) P.surf_dir[s][c] = s < p->n_surf ? p->surf_dir[s][c] : 0.0f;
  }
  const uint32_t hx = (uint32_t)(p->nx + 1) / 2u, hy = (uint32_t)(p->ny + 1) / 2u;
  P.div_hx = mcd_div_make(hx);
  P.div_ny = mcd_div_make((uint32_t)p->ny);
  P.div_nj = mcd_div_make(hy);
  P.div_ninj = mcd_div_make(hx * hy);
  P.seed_lo = (uint32_t)(p->seed & 0xFFFFFFFFu);
  P.seed_hi = (uint32_t)(p->seed >> 32);
  P.stream = p->stream;
  return nullptr;
}

extern "C" {

const char *mc_last_error(void) { return g_err.c_str(); }

int mc_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  int good = 0;
  for (int d = 0; d < n; ++d) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, d) != hipSuccess) continue;
    if (strncmp(prop.gcnArchName, "gfx950", 6) == 0) ++good;
  }
  return good;
}

int mc_create(const mc_params *p, const float *ni_aos, const int32_t *pt, const float *ns_aos, mc_handle **out) {
  if (!p || !ni_aos || !pt || !out) return fail(MC_EINVAL, "mc_create: null argument");
  *out = nullptr;
  if (p->nx < 1 || p->ny < 1 || p->nz < 1) return fail(MC_EINVAL, "mc_create: lattice dimensions must be >= 1");
  const long long ntl = (long long)p->nx * p->ny * p->nz;
  if (ntl >= (1LL << 31)) return fail(MC_EINVAL, "mc_create: lattice of 2^31 sites or more is not supported");
  const int dims[3] = {p->nx, p->ny, p->nz};
  for (int a = 0; a < 3; ++a) {
    if (p->bound[a] != MC_FREE && p->bound[a] != MC_PERIODIC)
      return fail(MC_EINVAL, "mc_create: boundary must be MC_FREE or MC_PERIODIC");
    // An odd periodic extent puts i=0 and i=N-1 (neighbours) on the same
    // sublattice and colour (reference src/evolve.cpp:105-113), so a pass
    // would update interacting sites concurrently.
    if (p->bound[a] == MC_PERIODIC && (dims[a] % 2) != 0)
      return fail(MC_EINVAL, "mc_create: a periodic axis needs an even number of sites");
  }
  if (p->potential < MC_LL || p->potential > MC_PEAR) return fail(MC_EINVAL, "mc_create: unknown potential");
  if (p->n_surf < 0 || p->n_surf > MC_MAX_SURF) return fail(MC_EINVAL, "mc_create: n_surf out of range");
  for (int s = 0; s < p->n_surf; ++s) {
    if (p->surf_kind[s] < MC_HOMEOTROPIC || p->surf_kind[s] > MC_STRONG_GHRL)
      return fail(MC_EINVAL, "mc_create: unknown anchoring kind");
    const int k = p->surf_kind[s];
    if (k == MC_RP || k == MC_STRONG || k == MC_RP_GHRL || k == MC_STRONG_GHRL) {
      const float *d = p->surf_dir[s];
      const float d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      if (!isfinite(d2) || fabsf(d2 - 1.0f) > 1e-3f)
        return fail(MC_EINVAL, "mc_create: surf_dir of surface " + std::to_string(s) +
                                   " must be the unit vector of (theta_s, phi_s), not the angles");
    }
  }
  if (p->order != MC_ORDER_CHECKER && p->order != MC_ORDER_REF8)
    return fail(MC_EINVAL, "mc_create: unknown sweep order");
  if (!(p->T > 0.0f) || !isfinite(p->T)) return fail(MC_EINVAL, "mc_create: temperature must be finite and > 0");
  const int nt = (int)ntl;
  bool all_bulk = true, any_surface = false;
  int valid = 0;
  for (int s = 0; s < nt; ++s) {
    if (pt[s] < 0) return fail(MC_EINVAL, "mc_create: negative point type");
    if (pt[s] > 1) {
      if (pt[s] - 2 >= p->n_surf) return fail(MC_EINVAL, "mc_create: point type names an undefined surface");
      any_surface = true;
    }
    if (pt[s] > 255) return fail(MC_EINVAL, "mc_create: point type above 255");
    if (pt[s] != 1) all_bulk = false;
    if (pt[s]) {
      ++valid;
      const float x = ni_aos[3 * (size_t)s], y = ni_aos[3 * (size_t)s + 1], z = ni_aos[3 * (size_t)s + 2];
      const float n2 = x * x + y * y + z * z;
      if (!isfinite(n2) || fabsf(n2 - 1.0f) > 1e-3f)
        return fail(MC_EINVAL, "mc_create: director at site " + std::to_string(s) + " is not a finite unit vector");
    }
  }
  if (valid == 0) return fail(MC_EINVAL, "mc_create: no active site");
  if (any_surface && !ns_aos) return fail(MC_EINVAL, "mc_create: surface sites need surface normals");
  // Slab geometry for the fast kernels: x, y periodic, z free, every site
  // active, and either no surface sites or exactly the faces of
  // src/geometry_slab.cpp:47-67 (pt 2 with normal (0,0,-1) on k = 0, pt 3
  // with (0,0,+1) on k = Nz-1).
  int zfree = 0, zsurf = 0;
  if (p->bound[0] == MC_PERIODIC && p->bound[1] == MC_PERIODIC && p->bound[2] != MC_PERIODIC && p->nz >= 2 &&
      p->nx % 4 == 0) {
    const int plane = p->nx * p->ny;
    bool faces = true;
    for (int s = 0; s < nt && (all_bulk || faces); ++s) {
      const int k = s / plane;
      const int want = k == 0 ? 2 : (k == p->nz - 1 ? 3 : 1);
      if (pt[s] != want) {
        faces = false;
        continue;
      }
      if (want > 1) {
        const float *n = ns_aos + 3 * (size_t)s;
        if (n[0] != 0.0f || n[1] != 0.0f || n[2] != (want == 2 ? -1.0f : 1.0f)) faces = false;
      }
    }
    if (all_bulk) zfree = 1;
    else if (faces) zfree = zsurf = 1;
  }
  if (mc_device_count() < 1) return fail(MC_ENODEV, "mc_create: no gfx950 device visible");

  mc_handle *h = new mc_handle();
  int rc = MC_OK;
  auto bail = [&](int code) {
    free_handle(h);
    return code;
  };
#define MC_HIPC(call)                                                                       \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess) {                                                                 \
      (void)hipGetLastError();                                                              \
      rc = fail(e_ == hipErrorOutOfMemory ? MC_ENOMEM : MC_EHIP,                            \
                std::string(#call) + " failed: " + hipGetErrorString(e_));                 \
      return bail(rc);                                                                      \
    }                                                                                       \
  } while (0)

  MC_HIPC(hipGetDevice(&h->device));
  hipDeviceProp_t prop;
  MC_HIPC(hipGetDeviceProperties(&prop, h->device));
  if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    rc = fail(MC_ENODEV, std::string("mc_create: current device is ") + prop.gcnArchName + ", not gfx950");
    return bail(rc);
  }
  h->cu_count = prop.multiProcessorCount;
  {
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, h->device) == hipSuccess && lds > 0)
      h->lds_limit = (size_t)lds;
    (void)hipGetLastError();
  }
  MC_HIPC(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  MC_HIPC(hipEventCreate(&h->ev0));
  MC_HIPC(hipEventCreate(&h->ev1));

  mc_to_dev_params(p, &h->P);
  h->P.zfree = zfree;
  h->P.zsurf = zsurf;
  h->P.nt = nt;
  h->P.valid = valid;
  h->order = p->order;

  const size_t fbytes = (size_t)nt * sizeof(float);
  MC_HIPC(hipMalloc(&h->nx, fbytes));
  MC_HIPC(hipMalloc(&h->ny, fbytes));
  MC_HIPC(hipMalloc(&h->nz, fbytes));
  MC_HIPC(hipMalloc(&h->staging, 3 * fbytes));
  MC_HIPC(hipMalloc(&h->st, sizeof(MCDevState)));
  MC_HIPC(hipMalloc(&h->ctr, (size_t)MCD_NSUB * MCD_SUB_STRIDE * sizeof(unsigned long long)));
  h->P.ctr = h->ctr;
  h->measure_grid = h->grid_for(nt);
  MC_HIPC(hipMalloc(&h->partials, (size_t)h->measure_grid * 8 * sizeof(double)));
  MC_HIPC(hipMalloc(&h->result, 16 * sizeof(double)));
  const int cgrid = h->grid_for(nt);
  MC_HIPC(hipMemcpyAsync(h->staging, ni_aos, 3 * fbytes, hipMemcpyHostToDevice, h->stream));
  mc_launch_aos_to_soa(h->stream, h->staging, h->nx, h->ny, h->nz, nt, cgrid);
  if (!all_bulk) {
    std::vector<uint8_t> pt8(nt);
    for (int s = 0; s < nt; ++s) pt8[s] = (uint8_t)pt[s];
    MC_HIPC(hipMalloc(&h->pt, (size_t)nt));
    MC_HIPC(hipMemcpyAsync(h->pt, pt8.data(), (size_t)nt, hipMemcpyHostToDevice, h->stream));
    MC_HIPC(hipStreamSynchronize(h->stream));
  }
  if (any_surface) {
    MC_HIPC(hipMalloc(&h->sx, fbytes));
    MC_HIPC(hipMalloc(&h->sy, fbytes));
    MC_HIPC(hipMalloc(&h->sz, fbytes));
    float *tmp = nullptr;
    MC_HIPC(hipMalloc(&tmp, 3 * fbytes));
    hipError_t e1 = hipMemcpyAsync(tmp, ns_aos, 3 * fbytes, hipMemcpyHostToDevice, h->stream);
    if (e1 == hipSuccess) mc_launch_aos_to_soa(h->stream, tmp, h->sx, h->sy, h->sz, nt, cgrid);
    hipError_t e2 = hipStreamSynchronize(h->stream);
    (void)hipFree(tmp);
    MC_HIPC(e1);
    MC_HIPC(e2);
  }
  // Measurement streams z-planes through LDS wherever the x4 layout applies
  // and a tile shape fits; the segment measurement covers the rest.
  if (mc_x4_eligible(h->launch()) &&
      mc_measure_slab_shape(h->P, h->lds_limit, h->cu_count, h->measure_grid, &h->mshape)) {
    h->measure_slab = true;
    (void)mc_slab_prepare(h->lds_limit);
  }
  k_init_state<<<1, 64, 0, h->stream>>>(h->st, p->T, p->elecE, p->ang_var, h->ctr);
  MC_HIPC(hipGetLastError());
  MC_HIPC(hipStreamSynchronize(h->stream));
  rc = select_kernel(h, MC_KERNEL_AUTO);
  if (rc) return bail(rc);
#undef MC_HIPC
  *out = h;
  return MC_OK;
}

int mc_destroy(mc_handle *h) {
  if (!h) return MC_OK;
  (void)hipSetDevice(h->device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  free_handle(h);
  return MC_OK;
}

static int set_state(mc_handle *h, int field, float v) {
  int rc = check_handle(h);
  if (rc) return rc;
  k_set_state_f<<<1, 1, 0, h->stream>>>(h->st, field, v);
  MC_LAUNCHED("set state");
  return MC_OK;
}

int mc_set_T(mc_handle *h, float T) {
  if (!(T > 0.0f) || !isfinite(T)) return fail(MC_EINVAL, "mc_set_T: temperature must be finite and > 0");
  return set_state(h, 0, T);
}
int mc_set_field(mc_handle *h, float elecE) {
  if (!isfinite(elecE)) return fail(MC_EINVAL, "mc_set_field: field must be finite");
  return set_state(h, 1, elecE);
}
int mc_set_ang_var(mc_handle *h, float a) {
  if (!isfinite(a) || a < 0.0f) return fail(MC_EINVAL, "mc_set_ang_var: ang_var must be finite and >= 0");
  return set_state(h, 2, a);
}

int mc_get_ang_var(mc_handle *h, float *a) {
  int rc = check_handle(h);
  if (rc) return rc;
  if (!a) return fail(MC_EINVAL, "mc_get_ang_var: null output");
  MCDevState s;
  MC_HIP(hipMemcpyAsync(&s, h->st, sizeof(s), hipMemcpyDeviceToHost, h->stream));
  MC_HIP(hipStreamSynchronize(h->stream));
  *a = s.ang_var;
  return MC_OK;
}

int mc_set_sweep_index(mc_handle *h, uint32_t sweep) {
  int rc = check_handle(h);
  if (rc) return rc;
  k_set_sweep_index<<<1, 1, 0, h->stream>>>(h->st, sweep);
  MC_LAUNCHED("set sweep index");
  return MC_OK;
}

int mc_get_acceptance(mc_handle *h, uint32_t *accepted, uint32_t *sweeps_done) {
  int rc = check_handle(h);
  if (rc) return rc;
  MCDevState s;
  MC_HIP(hipMemcpyAsync(&s, h->st, sizeof(s), hipMemcpyDeviceToHost, h->stream));
  MC_HIP(hipStreamSynchronize(h->stream));
  if (accepted) *accepted = s.last_acc;
  if (sweeps_done) *sweeps_done = s.sweep;
  return MC_OK;
}

int mc_set_sweep_kernel(mc_handle *h, int kind) {
  int rc = check_handle(h);
  if (rc) return rc;
  if (kind != MC_KERNEL_AUTO && kind != MC_KERNEL_GENERIC && kind != MC_KERNEL_X4 && kind != MC_KERNEL_SLAB)
    return fail(MC_EINVAL, "mc_set_sweep_kernel: unknown kernel kind");
  if ((kind == MC_KERNEL_X4 || kind == MC_KERNEL_SLAB) && !fast_eligible(h))
    return fail(MC_EINVAL, "mc_set_sweep_kernel: the segment and slab kernels need a checkerboard order and an "
                           "all-active lattice, periodic on every axis, with Nx a multiple of 4");
  MC_HIP(hipStreamSynchronize(h->stream));
  drop_graphs(h);
  return select_kernel(h, kind);
}

int mc_get_sweep_kernel(mc_handle *h, int *kind) {
  int rc = check_handle(h);
  if (rc) return rc;
  if (!kind) return fail(MC_EINVAL, "mc_get_sweep_kernel: null output");
  *kind = h->active;
  return MC_OK;
}

int mc_set_slab_tile(mc_handle *h, int ty, int cz, int xcd) {
  int rc = check_handle(h);
  if (rc) return rc;
  if (xcd != 0 && xcd != 1) return fail(MC_EINVAL, "mc_set_slab_tile: xcd must be 0 or 1");
  MCSlabShape s{0, 0, xcd};
  if (ty != 0 || cz != 0) {
    s = MCSlabShape{ty, cz, xcd};
    if (!mc_slab_shape_ok(h->P, s, h->lds_limit))
      return fail(MC_EINVAL, "mc_set_slab_tile: ty must divide Ny and cz divide Nz, with (Nx/4)*(ty+4) <= 768 "
                             "threads and the tile's LDS within the device limit");
  }
  MC_HIP(hipStreamSynchronize(h->stream));
  drop_graphs(h);
  h->slab = s;
  return MC_OK;
}

int mc_get_slab_tile(mc_handle *h, int *ty, int *cz, int *xcd) {
  int rc = check_handle(h);
  if (rc) return rc;
  const MCSlabShape s = slab_shape(h);
  if (ty) *ty = s.ty;
  if (cz) *cz = s.cz;
  if (xcd) *xcd = s.xcd;
  return MC_OK;
}

int mc_set_tuning(mc_handle *h, int grid, int xcd) {
  int rc = check_handle(h);
  if (rc) return rc;
  if (grid < 0 || (xcd && grid % 8) || (xcd != 0 && xcd != 1))
    return fail(MC_EINVAL, "mc_set_tuning: grid must be >= 0 (0 = default) and a multiple of 8 with xcd = 1");
  MC_HIP(hipStreamSynchronize(h->stream));
  drop_graphs(h);
  h->x4_grid = grid;
  h->x4_xcd = xcd;
  return MC_OK;
}

static int autotune_slab(mc_handle *h, int *grid, int *xcd, float *best_us) {
  // Candidates sweep a scratch copy of the lattice and state (two plane sets,
  // swapped after each sweep), so tuning leaves the chain untouched.
  const size_t nt = (size_t)h->P.nt;
  const MCDevParams &P = h->P;
  float *scratch = nullptr;
  MCDevState *sst = nullptr;
  MC_HIP(hipMalloc(&scratch, 6 * nt * sizeof(float)));
  if (hipMalloc(&sst, sizeof(MCDevState)) != hipSuccess) {
    (void)hipFree(scratch);
    return fail(MC_ENOMEM, "mc_autotune: scratch state allocation failed");
  }
  MCSlabShape best_s = slab_shape(h);
  float best = 1e30f;
  hipError_t err = hipSuccess;
  const int tys[] = {1, 2, 4, 6, 8, 10, 12, 16, 20, 24, 32};
  for (int ty : tys) {
    for (int cz = P.nz; cz >= 1 && err == hipSuccess; cz = (cz % 2 == 0) ? cz / 2 : 0) {
      if (cz < 4 && P.nz >= 8) break;
      for (int xm = 0; xm < 2 && err == hipSuccess; ++xm) {
        const MCSlabShape cand{ty, cz, xm};
        if (!mc_slab_shape_ok(P, cand, h->lds_limit)) continue;
        if (xm && ((P.ny / ty) * (P.nz / cz)) % 8) continue;
        float *io[2] = {scratch, scratch + 3 * nt};
        for (int d = 0; d < 3 && err == hipSuccess; ++d) {
          float *src = d == 0 ? h->nx : (d == 1 ? h->ny : h->nz);
          err = hipMemcpyAsync(io[0] + d * nt, src, nt * sizeof(float), hipMemcpyDeviceToDevice, h->stream);
        }
        if (err == hipSuccess) err = hipMemcpyAsync(sst, h->st, sizeof(MCDevState), hipMemcpyDeviceToDevice, h->stream);
        if (err != hipSuccess) break;
        MCLaunch L = h->launch();
        L.st = sst;
        int cur = 0;
        auto sweep = [&]() {
          L.nx = io[cur];
          L.ny = io[cur] + nt;
          L.nz = io[cur] + 2 * nt;
          float *o = io[cur ^ 1];
          mc_launch_sweep_slab(L, o, o + nt, o + 2 * nt, cand, nullptr);
          cur ^= 1;
        };
        sweep();  // warm-up
        sweep();
        const int reps = 6;
        err = hipEventRecord(h->ev0, h->stream);
        for (int w = 0; w < reps; ++w) sweep();
        if (err == hipSuccess) err = hipGetLastError();
        if (err == hipSuccess) err = hipEventRecord(h->ev1, h->stream);
        if (err == hipSuccess) err = hipEventSynchronize(h->ev1);
        float ms = 0.0f;
        if (err == hipSuccess) err = hipEventElapsedTime(&ms, h->ev0, h->ev1);
        if (err != hipSuccess) break;
        const float us = ms * 1e3f / reps;
        if (us < best) {
          best = us;
          best_s = cand;
        }
      }
    }
  }
  (void)hipFree(scratch);
  (void)hipFree(sst);
  MC_HIP(err);
  int rc = mc_set_slab_tile(h, best_s.ty, best_s.cz, best_s.xcd);
  if (rc) return rc;
  if (grid) *grid = (P.ny / best_s.ty) * (P.nz / best_s.cz);
  if (xcd) *xcd = best_s.xcd;
  if (best_us) *best_us = best;
  return MC_OK;
}

static int autotune_x4(mc_handle *h, int *grid, int *xcd, float *best_us);

int mc_autotune(mc_handle *h, int *grid, int *xcd, float *best_us) {
  int rc = check_handle(h);
  if (rc) return rc;
  if (h->kernel == MC_KERNEL_AUTO && (use_slab(h) || use_x4(h))) {
    // Time both fast kernels at their best shapes and keep the faster.
    int gs = 0, xs = 0, gx = 0, xx = 0;
    float us_s = 1e30f, us_x = 1e30f;
    bool have_slab = select_kernel(h, MC_KERNEL_SLAB) == MC_OK;
    if (have_slab) {
      rc = autotune_slab(h, &gs, &xs, &us_s);
      if (rc) return rc;
    }
    rc = select_kernel(h, MC_KERNEL_X4);
    if (rc) return rc;
    rc = autotune_x4(h, &gx, &xx, &us_x);
    if (rc) return rc;
    const bool slab_wins = have_slab && us_s <= us_x;
    if (slab_wins) {
      rc = select_kernel(h, MC_KERNEL_SLAB);
      if (rc) return rc;
    }
    h->kernel = MC_KERNEL_AUTO;
    MC_HIP(hipStreamSynchronize(h->stream));
    drop_graphs(h);
    if (grid) *grid = slab_wins ? gs : gx;
    if (xcd) *xcd = slab_wins ? xs : xx;
    if (best_us) *best_us = slab_wins ? us_s : us_x;
    return MC_OK;
  }
  if (use_slab(h)) return autotune_slab(h, grid, xcd, best_us);
  if (!use_x4(h)) {
    if (grid) *grid = 0;
    if (xcd) *xcd = 0;
    return MC_OK;
  }
  return autotune_x4(h, grid, xcd, best_us);
}

static int autotune_x4(mc_handle *h, int *grid, int *xcd, float *best_us) {
  // Candidates run on a scratch copy of the lattice and state, so tuning
  // leaves the chain untouched.
  int rc = MC_OK;
  const size_t nt = (size_t)h->P.nt;
  float *scratch = nullptr;
  MCDevState *sst = nullptr;
  MC_HIP(hipMalloc(&scratch, 3 * nt * sizeof(float)));
  if (hipMalloc(&sst, sizeof(MCDevState)) != hipSuccess) {
    (void)hipFree(scratch);
    return fail(MC_ENOMEM, "mc_autotune: scratch state allocation failed");
  }
  MCLaunch L = h->launch();
  L.nx = scratch;
  L.ny = scratch + nt;
  L.nz = scratch + 2 * nt;
  L.st = sst;
  int best_g = 0, best_x = 0;
  float best = 1e30f;
  const int mults[] = {2, 4, 8, 16, 32};
  hipError_t err = hipSuccess;
  for (int xm = 0; xm < 2 && err == hipSuccess; ++xm) {
    for (int m : mults) {
      const int g = h->cu_count * m;
      if (xm && g % 8) continue;
      err = hipMemcpyAsync(scratch, h->nx, nt * sizeof(float), hipMemcpyDeviceToDevice, h->stream);
      if (err == hipSuccess) err = hipMemcpyAsync(scratch + nt, h->ny, nt * sizeof(float), hipMemcpyDeviceToDevice, h->stream);
      if (err == hipSuccess) err = hipMemcpyAsync(scratch + 2 * nt, h->nz, nt * sizeof(float), hipMemcpyDeviceToDevice, h->stream);
      if (err == hipSuccess) err = hipMemcpyAsync(sst, h->st, sizeof(MCDevState), hipMemcpyDeviceToDevice, h->stream);
      if (err != hipSuccess) break;
      for (int w = 0; w < 2; ++w) {  // warm-up
        mc_launch_sweep_x4(L, 0, g, 0, xm, nullptr);
        mc_launch_sweep_x4(L, 1, g, 1, xm, nullptr);
      }
      const int reps = 6;
      err = hipEventRecord(h->ev0, h->stream);
      for (int w = 0; w < reps; ++w) {
        mc_launch_sweep_x4(L, 0, g, 0, xm, nullptr);
        mc_launch_sweep_x4(L, 1, g, 1, xm, nullptr);
      }
      if (err == hipSuccess) err = hipEventRecord(h->ev1, h->stream);
      if (err == hipSuccess) err = hipEventSynchronize(h->ev1);
      float ms = 0.0f;
      if (err == hipSuccess) err = hipEventElapsedTime(&ms, h->ev0, h->ev1);
      if (err != hipSuccess) break;
      const float us = ms * 1e3f / reps;
      if (us < best) {
        best = us;
        best_g = g;
        best_x = xm;
      }
    }
  }
  (void)hipFree(scratch);
  (void)hipFree(sst);
  MC_HIP(err);
  rc = mc_set_tuning(h, best_g, best_x);
  if (rc) return rc;
  if (grid) *grid = best_g;
  if (xcd) *xcd = best_x;
  if (best_us) *best_us = best;
  return MC_OK;
}

int mc_sweeps(mc_handle *h, int n) {
  int rc = check_handle(h);
  if (rc) return rc;
  if (n < 0) return fail(MC_EINVAL, "mc_sweeps: negative sweep count");
  for (int w = 0; w < n; ++w) {
    rc = enqueue_sweep(h, nullptr);
    if (rc) return rc;
  }
  return MC_OK;
}

int mc_sweeps_graph(mc_handle *h, int n) {
  int rc = check_handle(h);
  if (rc) return rc;
  if (n < 0) return fail(MC_EINVAL, "mc_sweeps_graph: negative sweep count");
  if (n == 0) return MC_OK;
  if (!graphable(h)) return mc_sweeps(h, n);
  GraphPlan g;
  rc = plan_sweep_graphs(h, n, &g);
  if (rc) return rc;
  return launch_sweep_graphs(h, g);
}

int mc_sync(mc_handle *h) {
  int rc = check_handle(h);
  if (rc) return rc;
  MC_HIP(hipStreamSynchronize(h->stream));
  return MC_OK;
}

int mc_time_sweeps(mc_handle *h, int n, int use_graph, float *ms, int *launches) {
  int rc = check_handle(h);
  if (rc) return rc;
  if (n < 1 || !ms) return fail(MC_EINVAL, "mc_time_sweeps: need n >= 1 and an output");
  if (use_graph && !graphable(h)) use_graph = 0;
  GraphPlan g;
  if (use_graph) {
    // Instantiate before timing so capture cost stays outside the region.
    MC_HIP(hipStreamSynchronize(h->stream));
    rc = plan_sweep_graphs(h, n, &g);
    if (rc) return rc;
  }
  const int per = passes_per_sweep(h);
  MC_HIP(hipEventRecord(h->ev0, h->stream));
  rc = use_graph ? launch_sweep_graphs(h, g) : mc_sweeps(h, n);
  if (rc) return rc;
  MC_HIP(hipEventRecord(h->ev1, h->stream));
  MC_HIP(hipEventSynchronize(h->ev1));
  MC_HIP(hipEventElapsedTime(ms, h->ev0, h->ev1));
  if (launches) *launches = n * per;
  return MC_OK;
}

int mc_time_passes(mc_handle *h, int n, int every, double *avg_pass_ms, double *total_ms, int *passes) {
  int rc = check_handle(h);
  if (rc) return rc;
  if (n < 1 || every < 0 || !avg_pass_ms)
    return fail(MC_EINVAL, "mc_time_passes: need n >= 1, every >= 0 and an output");
  int np = 0;
  const long long sweeps0 = h->host_sweeps;
  for (int w = 0; w < n; ++w, ++h->host_sweeps) np += passes_per_sweep(h);  // the loop below, dry
  h->host_sweeps = sweeps0;
  const int nt = every ? (np + every - 1) / every : 0;  // bracketed passes: q = 0, every, 2 every, ...
  std::vector<hipEvent_t> ev(2 * (size_t)nt, nullptr);
  for (auto &e : ev) MC_HIP(hipEventCreate(&e));
  MC_HIP(hipEventRecord(h->ev0, h->stream));
  for (int w = 0, q = 0; w < n; ++w) {
    const int per = passes_per_sweep(h);
    for (int t = 0; t < per; ++t, ++q) {
      const bool timed = every && q % every == 0;
      if (timed) MC_HIP(hipEventRecord(ev[2 * (q / every)], h->stream));
      launch_pass(h, t, nullptr);
      if (timed) MC_HIP(hipEventRecord(ev[2 * (q / every) + 1], h->stream));
    }
    ++h->host_sweeps;
  }
  MC_LAUNCHED("timed sweep pass");
  MC_HIP(hipEventRecord(h->ev1, h->stream));
  MC_HIP(hipEventSynchronize(h->ev1));
  double sum = 0.0;
  for (int i = 0; i < nt; ++i) {
    float ms = 0.0f;
    MC_HIP(hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]));
    sum += ms;
  }
  float tot = 0.0f;
  MC_HIP(hipEventElapsedTime(&tot, h->ev0, h->ev1));
  for (auto &e : ev) (void)hipEventDestroy(e);
  *avg_pass_ms = nt ? sum / nt : (double)tot / np;
  if (total_ms) *total_ms = tot;
  if (passes) *passes = np;
  return MC_OK;
}

int mc_measure(mc_handle *h, double *E, double *S, double *q6) {
  int rc = check_handle(h);
  if (rc) return rc;
  launch_measure(h, nullptr);
  MC_LAUNCHED("measure");
  double r[9];
  MC_HIP(hipMemcpyAsync(r, h->result, sizeof(r), hipMemcpyDeviceToHost, h->stream));
  MC_HIP(hipStreamSynchronize(h->stream));
  if (E) *E = r[0];
  if (S) *S = r[1];
  if (q6)
    for (int c = 0; c < 6; ++c) q6[c] = r[2 + c];
  return MC_OK;
}

int mc_measure_series(mc_handle *h, int n, double *E_series, double *S_series) {
  return mc_measure_series_every(h, n, 1, E_series, S_series);
}

int mc_measure_series_every(mc_handle *h, int n, int every, double *E_series, double *S_series) {
  int rc = check_handle(h);
  if (rc) return rc;
  if (n < 1 || every < 1) return fail(MC_EINVAL, "mc_measure_series: need n >= 1 and every >= 1");
  if (n > h->series_cap) {
    if (h->series) MC_HIP(hipFree(h->series));
    h->series = nullptr;
    h->series_cap = 0;
    MC_HIP(hipMalloc(&h->series, 2 * (size_t)n * sizeof(double)));
    h->series_cap = n;
  }
  for (int w = 0; w < n; ++w) {
    for (int e = 0; e < every; ++e) {
      rc = enqueue_sweep(h, nullptr);
      if (rc) return rc;
    }
    // (re-read after the sweep: the slab kernel moves the lattice to the other plane set)
    launch_measure(h, h->series + 2 * (size_t)w);
  }
  MC_LAUNCHED("measured sweep");
  std::vector<double> buf(2 * (size_t)n);
  MC_HIP(hipMemcpyAsync(buf.data(), h->series, buf.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  MC_HIP(hipStreamSynchronize(h->stream));
  for (int w = 0; w < n; ++w) {
    if (E_series) E_series[w] = buf[2 * (size_t)w];
    if (S_series) S_series[w] = buf[2 * (size_t)w + 1];
  }
  return MC_OK;
}

int mc_site_energies(mc_handle *h, float *out) {
  int rc = check_handle(h);
  if (rc) return rc;
  if (!out) return fail(MC_EINVAL, "mc_site_energies: null output");
  rc = ensure_parity_buffers(h);
  if (rc) return rc;
  mc_launch_site_energy(h->launch(), h->grid_for(h->P.nt), h->pbuf);
  MC_LAUNCHED("site energy");
  MC_HIP(hipMemcpyAsync(out, h->pbuf, (size_t)h->P.nt * sizeof(float), hipMemcpyDeviceToHost, h->stream));
  MC_HIP(hipStreamSynchronize(h->stream));
  return MC_OK;
}

int mc_local_order(mc_handle *h, float *S_out) {
  int rc = check_handle(h);
  if (rc) return rc;
  if (!S_out) return fail(MC_EINVAL, "mc_local_order: null output");
  rc = ensure_parity_buffers(h);
  if (rc) return rc;
  mc_launch_local_S(h->launch(), h->grid_for(h->P.nt), h->pbuf);
  MC_LAUNCHED("local order parameter");
  MC_HIP(hipMemcpyAsync(S_out, h->pbuf, (size_t)h->P.nt * sizeof(float), hipMemcpyDeviceToHost, h->stream));
  MC_HIP(hipStreamSynchronize(h->stream));
  return MC_OK;
}

int mc_get_directors_aos(mc_handle *h, float *ni_aos) {
  int rc = check_handle(h);
  if (rc) return rc;
  if (!ni_aos) return fail(MC_EINVAL, "mc_get_directors_aos: null output");
  const int nt = h->P.nt;
  mc_launch_soa_to_aos(h->stream, h->nx, h->ny, h->nz, h->staging, nt, h->grid_for(nt));
  MC_LAUNCHED("soa_to_aos");
  MC_HIP(hipMemcpyAsync(ni_aos, h->staging, 3 * (size_t)nt * sizeof(float), hipMemcpyDeviceToHost, h->stream));
  MC_HIP(hipStreamSynchronize(h->stream));
  return MC_OK;
}

int mc_set_directors_aos(mc_handle *h, const float *ni_aos) {
  int rc = check_handle(h);
  if (rc) return rc;
  if (!ni_aos) return fail(MC_EINVAL, "mc_set_directors_aos: null input");
  const int nt = h->P.nt;
  MC_HIP(hipMemcpyAsync(h->staging, ni_aos, 3 * (size_t)nt * sizeof(float), hipMemcpyHostToDevice, h->stream));
  mc_launch_aos_to_soa(h->stream, h->staging, h->nx, h->ny, h->nz, nt, h->grid_for(nt));
  MC_LAUNCHED("aos_to_soa");
  MC_HIP(hipStreamSynchronize(h->stream));
  return MC_OK;
}

int mc_pass(mc_handle *h, int set, const float *prop_aos, const float *u, uint8_t *accept, float *e_old,
            float *e_new) {
  int rc = check_handle(h);
  if (rc) return rc;
  if (set < 0 || set > 9) return fail(MC_EINVAL, "mc_pass: set must be 0..7 (sublattice) or 8/9 (colour)");
  if (!prop_aos || !u) return fail(MC_EINVAL, "mc_pass: proposals and uniforms are required");
  rc = ensure_parity_buffers(h);
  if (rc) return rc;
  const size_t nt = (size_t)h->P.nt;
  float *d_u = h->pbuf, *d_eo = h->pbuf + nt, *d_en = h->pbuf + 2 * nt;
  MC_HIP(hipMemcpyAsync(h->staging, prop_aos, 3 * nt * sizeof(float), hipMemcpyHostToDevice, h->stream));
  MC_HIP(hipMemcpyAsync(d_u, u, nt * sizeof(float), hipMemcpyHostToDevice, h->stream));
  MC_HIP(hipMemsetAsync(d_eo, 0, 2 * nt * sizeof(float), h->stream));
  MC_HIP(hipMemsetAsync(h->abuf, 0, nt, h->stream));
  mc_launch_pass_inject(h->launch(), set, h->grid_for(mc_set_slots(h->P, set)), h->staging, d_u, h->abuf, d_eo,
                        d_en);
  MC_LAUNCHED("pass_inject");
  if (accept) MC_HIP(hipMemcpyAsync(accept, h->abuf, nt, hipMemcpyDeviceToHost, h->stream));
  if (e_old) MC_HIP(hipMemcpyAsync(e_old, d_eo, nt * sizeof(float), hipMemcpyDeviceToHost, h->stream));
  if (e_new) MC_HIP(hipMemcpyAsync(e_new, d_en, nt * sizeof(float), hipMemcpyDeviceToHost, h->stream));
  MC_HIP(hipStreamSynchronize(h->stream));
  return MC_OK;
}

int mc_record_sweep(mc_handle *h, float *u0, float *u1, float *u2, float *prop_aos, uint8_t *accept, float *e_old,
                    float *e_new) {
  int rc = check_handle(h);
  if (rc) return rc;
  rc = ensure_parity_buffers(h);
  if (rc) return rc;
  const size_t nt = (size_t)h->P.nt;
  // Planes: u0 u1 u2 eo en in pbuf, proposals in staging (SoA thirds).
  MCRecord rec;
  rec.u0 = h->pbuf;
  rec.u1 = h->pbuf + nt;
  rec.u2 = h->pbuf + 2 * nt;
  rec.eo = h->pbuf + 3 * nt;
  rec.en = h->pbuf + 4 * nt;
  rec.px = h->staging;
  rec.py = h->staging + nt;
  rec.pz = h->staging + 2 * nt;
  rec.acc = h->abuf;
  MC_HIP(hipMemsetAsync(h->pbuf, 0, 5 * nt * sizeof(float), h->stream));
  MC_HIP(hipMemsetAsync(h->staging, 0, 3 * nt * sizeof(float), h->stream));
  MC_HIP(hipMemsetAsync(h->abuf, 0, nt, h->stream));
  rc = enqueue_sweep(h, &rec);
  if (rc) return rc;
  if (u0) MC_HIP(hipMemcpyAsync(u0, rec.u0, nt * sizeof(float), hipMemcpyDeviceToHost, h->stream));
  if (u1) MC_HIP(hipMemcpyAsync(u1, rec.u1, nt * sizeof(float), hipMemcpyDeviceToHost, h->stream));
  if (u2) MC_HIP(hipMemcpyAsync(u2, rec.u2, nt * sizeof(float), hipMemcpyDeviceToHost, h->stream));
  if (e_old) MC_HIP(hipMemcpyAsync(e_old, rec.eo, nt * sizeof(float), hipMemcpyDeviceToHost, h->stream));
  if (e_new) MC_HIP(hipMemcpyAsync(e_new, rec.en, nt * sizeof(float), hipMemcpyDeviceToHost, h->stream));
  if (accept) MC_HIP(hipMemcpyAsync(accept, rec.acc, nt, hipMemcpyDeviceToHost, h->stream));
  if (prop_aos) {
    // Interleave the SoA proposal planes into the caller's AoS array via the
    // pbuf tail (2 spare planes are not enough for 3, so go through host).
    std::vector<float> soa(3 * nt);
    MC_HIP(hipMemcpyAsync(soa.data(), h->staging, 3 * nt * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    MC_HIP(hipStreamSynchronize(h->stream));
    for (size_t s = 0; s < nt; ++s) {
      prop_aos[3 * s] = soa[s];
      prop_aos[3 * s + 1] = soa[nt + s];
      prop_aos[3 * s + 2] = soa[2 * nt + s];
    }
  }
  MC_HIP(hipStreamSynchronize(h->stream));
  return MC_OK;
}

int mc_set_learned_proposal(mc_handle *h, mc_mlp *net, int every) {
  int rc = check_handle(h);
  if (rc) return rc;
  if (every < 0) return fail(MC_EINVAL, "mc_set_learned_proposal: every must be >= 0");
  if (net && every > 0) {
    if (h->order != MC_ORDER_CHECKER || h->pt)
      return fail(MC_EINVAL, "mc_set_learned_proposal: the learned proposal needs the checkerboard order and an "
                             "all-active lattice");
    if (mc_mlp_device(net) != h->device)
      return fail(MC_EINVAL, "mc_set_learned_proposal: network and lattice live on different devices");
  }
  MC_HIP(hipStreamSynchronize(h->stream));
  drop_graphs(h);  // captured sweeps bake in the proposal
  h->learned = every > 0 ? net : nullptr;
  h->learned_every = net ? every : 0;
  h->host_sweeps = 0;
  return MC_OK;
}

int mc_set_learned_reflection(mc_handle *h, int on) {
  int rc = check_handle(h);
  if (rc) return rc;
  if (on != 0 && on != 1) return fail(MC_EINVAL, "mc_set_learned_reflection: on must be 0 or 1");
  MC_HIP(hipStreamSynchronize(h->stream));
  if (on != h->learned_reflect) drop_graphs(h);  // captured learned sweeps bake in the switch
  h->learned_reflect = on;
  return MC_OK;
}

int mc_record_learned_sweep(mc_handle *h, float *mu, float *logk, float *prop_aos, float *u, uint8_t *accept,
                            float *dE, float *refl_aos, float *u2, uint8_t *accept2, float *dE2) {
  int rc = check_handle(h);
  if (rc) return rc;
  if (!h->learned) return fail(MC_ESTATE, "mc_record_learned_sweep: no learned proposal attached");
  rc = ensure_parity_buffers(h);
  if (rc) return rc;
  const size_t nt = (size_t)h->P.nt;
  if (!h->lrbuf) MC_HIP(hipMalloc(&h->lrbuf, 5 * nt * sizeof(float)));
  if (!h->lrabuf) MC_HIP(hipMalloc(&h->lrabuf, nt));
  MCLearnRecord rec;
  rec.mu = h->staging;  // 3 planes, AoS
  rec.prop = h->pbuf;   // 3 planes, AoS
  rec.logk = h->pbuf + 3 * nt;
  rec.u = h->pbuf + 4 * nt;
  rec.dE = h->pbuf + 5 * nt;
  rec.acc = h->abuf;
  rec.refl = h->lrbuf;  // 3 planes, AoS
  rec.u2 = h->lrbuf + 3 * nt;
  rec.dE2 = h->lrbuf + 4 * nt;
  rec.acc2 = h->lrabuf;
  MC_HIP(hipMemsetAsync(h->pbuf, 0, 6 * nt * sizeof(float), h->stream));
  MC_HIP(hipMemsetAsync(h->staging, 0, 3 * nt * sizeof(float), h->stream));
  MC_HIP(hipMemsetAsync(h->abuf, 0, nt, h->stream));
  MC_HIP(hipMemsetAsync(h->lrbuf, 0, 5 * nt * sizeof(float), h->stream));
  MC_HIP(hipMemsetAsync(h->lrabuf, 0, nt, h->stream));
  const long long keep = h->host_sweeps;
  h->host_sweeps = 0;  // force a learned sweep
  const int every = h->learned_every;
  h->learned_every = 1;
  rc = enqueue_sweep(h, nullptr, &rec);
  h->learned_every = every;
  h->host_sweeps = keep + 1;
  if (rc) return rc;
  if (mu) MC_HIP(hipMemcpyAsync(mu, rec.mu, 3 * nt * sizeof(float), hipMemcpyDeviceToHost, h->stream));
  if (prop_aos) MC_HIP(hipMemcpyAsync(prop_aos, rec.prop, 3 * nt * sizeof(float), hipMemcpyDeviceToHost, h->stream));
  if (logk) MC_HIP(hipMemcpyAsync(logk, rec.logk, nt * sizeof(float), hipMemcpyDeviceToHost, h->stream));
  if (u) MC_HIP(hipMemcpyAsync(u, rec.u, nt * sizeof(float), hipMemcpyDeviceToHost, h->stream));
  if (dE) MC_HIP(hipMemcpyAsync(dE, rec.dE, nt * sizeof(float), hipMemcpyDeviceToHost, h->stream));
  if (accept) MC_HIP(hipMemcpyAsync(accept, rec.acc, nt, hipMemcpyDeviceToHost, h->stream));
  if (refl_aos) MC_HIP(hipMemcpyAsync(refl_aos, rec.refl, 3 * nt * sizeof(float), hipMemcpyDeviceToHost, h->stream));
  if (u2) MC_HIP(hipMemcpyAsync(u2, rec.u2, nt * sizeof(float), hipMemcpyDeviceToHost, h->stream));
  if (dE2) MC_HIP(hipMemcpyAsync(dE2, rec.dE2, nt * sizeof(float), hipMemcpyDeviceToHost, h->stream));
  if (accept2) MC_HIP(hipMemcpyAsync(accept2, rec.acc2, nt, hipMemcpyDeviceToHost, h->stream));
  MC_HIP(hipStreamSynchronize(h->stream));
  return MC_OK;
}

}  // extern "C"

int mc_handle_launch(mc_handle *h, MCLaunch *out) {
  int rc = check_handle(h);
  if (rc) return rc;
  *out = h->launch();
  return MC_OK;
}
