// Batched replicas: many independent bulk lattices of uneven sizes in one
// device allocation, swept by one persistent launch per pass and replayed as
// one captured graph (declared in include/mclics_hip.h).
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <string>
#include <vector>

#include "../../include/mclics_hip.h"
#include "mc_kernels.h"

namespace {

int bfail(int code, const std::string &msg) {
  mc_set_error_message(msg.c_str());
  return code;
}

#define MB_HIP(call)                                                                          \
  do {                                                                                        \
    hipError_t e_ = (call);                                                                   \
    if (e_ != hipSuccess) {                                                                   \
      (void)hipGetLastError();                                                                \
      return bfail(e_ == hipErrorOutOfMemory ? MC_ENOMEM : MC_EHIP,                           \
                   std::string(#call) + " failed: " + hipGetErrorString(e_));                \
    }                                                                                         \
  } while (0)

__global__ void k_batch_init(MCDevState *st, int nrep, const float *T, const float *E, const float *av) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrep) return;
  st[r].T = T[r];
  st[r].elecE = E[r];
  st[r].ang_var = av[r];
  st[r].sweep = 0u;
  st[r].acc_count = 0u;
  st[r].blocks_done = 0u;
  st[r].last_acc = 0u;
  st[r].meas_done = 0u;
}

__global__ void k_batch_set_T(MCDevState *st, int r, float T) { st[r].T = T; }

}  // namespace

struct mc_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  int nrep = 0;
  int potential = 0;
  int cu_count = 256;
  int grid = 0;
  int xcd = 0;
  std::vector<MCDevParams> hp;
  std::vector<long long> off;
  std::vector<int> tile0, ntile;
  int ntiles = 0;
  long long total_sites = 0;
  MCDevParams *d_params = nullptr;
  MCDevState *d_states = nullptr;
  int *d_tile_rep = nullptr, *d_rep_tile0 = nullptr, *d_rep_ntiles = nullptr;
  long long *d_rep_off = nullptr;
  float *nx = nullptr, *ny = nullptr, *nz = nullptr;
  uint32_t *d_done = nullptr;
  double *d_partials = nullptr, *d_out = nullptr, *d_series = nullptr;
  int series_cap = 0;
  float *staging = nullptr;
  std::map<int, hipGraphExec_t> graphs;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // Slab sweeps: a second plane set (flip tells which allocation holds the
  // current lattice), the workgroup and tile-shape tables of the current
  // shape, and the shape's parameters (thread budget, planes per tile).
  int kind = MC_KERNEL_AUTO;  // requested sweep kernel
  bool slab = false;          // active: slab (true) or segment (false)
  // Persistent slab sweeps (MC_KERNEL_SLAB_PERSISTENT): up to kPersistChunk
  // sweeps per launch of persist_grid workgroups, with the launch's barrier
  // and tile counters (and the fault flag, word 1) in d_sync.
  bool persist = false;
  int persist_grid = 0;
  unsigned int *d_sync = nullptr;
  unsigned int spin_limit = 1u << 21;
  float *bx = nullptr, *by = nullptr, *bz = nullptr;
  int flip = 0;
  int4 *d_wg = nullptr;
  int2 *d_geo = nullptr;
  MCBatchSlab S{};
  int slab_budget = 0, slab_cz = 0;
  bool slab_aligned = false;
  size_t lds_limit = 65536;
  // Measurement tiles (plane streaming, independent of the sweep kernel and
  // its shape): the workgroup table and per-replica shapes, each replica's
  // first workgroup and workgroup count, and the launch shape.  Empty when
  // some replica has no measurement tile (then segment tiles are used).
  int4 *d_mwg = nullptr;
  int2 *d_mgeo = nullptr;
  int *d_mwg0 = nullptr, *d_mnwg = nullptr;
  MCBatchSlab M{};
  int mthreads = 0;
  size_t mlds = 0;

  MCBatchLaunch launch() const {
    MCBatchLaunch B;
    B.stream = stream;
    B.potential = potential;
    B.nrep = nrep;
    B.ntiles = ntiles;
    B.params = d_params;
    B.states = d_states;
    B.tile_rep = d_tile_rep;
    B.rep_tile0 = d_rep_tile0;
    B.rep_ntiles = d_rep_ntiles;
    B.rep_off = d_rep_off;
    B.nx = nx;
    B.ny = ny;
    B.nz = nz;
    B.blocks_done = d_done;
    B.xcd = xcd;
    return B;
  }
  int sweep_grid() const { return grid > 0 ? grid : cu_count * 8; }
};

namespace {

void free_batch(mc_batch *b) {
  if (!b) return;
  for (auto &kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
  void *ptrs[] = {b->d_params, b->d_states, b->d_tile_rep, b->d_rep_tile0, b->d_rep_ntiles, b->d_rep_off, b->nx,
                  b->ny, b->nz, b->d_done, b->d_partials, b->d_out, b->d_series, b->staging, b->bx, b->by, b->bz,
                  b->d_wg, b->d_geo, b->d_mwg, b->d_mgeo, b->d_mwg0, b->d_mnwg, b->d_sync};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

int check_batch(mc_batch *b) {
  if (!b) return bfail(MC_EINVAL, "null batch");
  if (hipSetDevice(b->device) != hipSuccess) return bfail(MC_EHIP, "hipSetDevice failed for the batch's device");
  return MC_OK;
}

void swap_planes(mc_batch *b) {
  std::swap(b->nx, b->bx);
  std::swap(b->ny, b->by);
  std::swap(b->nz, b->bz);
  b->flip ^= 1;
}

void drop_graphs(mc_batch *b) {
  for (auto &kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
  b->graphs.clear();
}

// Tile shape of every replica for a thread budget and a planes-per-tile
// target: ty the largest divisor of Ny with (Nx/4)(ty+4) <= budget, cz the
// largest divisor of Nz <= cz_target, both checked as for one lattice.
// Fills the host tables; false when some replica has no such shape.
bool slab_plan(const mc_batch *b, int budget, int cz_target, bool aligned, std::vector<int4> &wg,
               std::vector<int2> &geo, int *threads, size_t *lds) {
  wg.clear();
  geo.assign(b->nrep, int2{0, 0});
  *threads = 0;
  *lds = 0;
  for (int r = 0; r < b->nrep; ++r) {
    const MCDevParams &P = b->hp[r];
    const int S = P.nx / 4;
    int ty = 0, cz = 0;
    // aligned: prefer a tile whose owned rows of each index parity fill
    // whole waves (S * ceil(ty / 2) a multiple of 64), so no wave of the
    // tile mixes the two colour layouts of a segment
    for (int pass = aligned ? 0 : 1; pass < 2 && ty == 0; ++pass)
      for (int t = P.ny; t >= 1; --t)
        if (P.ny % t == 0 && S * (t + 4) <= budget && (pass == 1 || (S * ((t + 1) / 2)) % 64 == 0)) {
          ty = t;
          break;
        }
    for (int c = std::min(cz_target, P.nz); c >= 1; --c)
      if (P.nz % c == 0) {
        cz = c;
        break;
      }
    const MCSlabShape sh{ty, cz, 0};
    if (ty < 1 || cz < 1 || !mc_slab_shape_ok(P, sh, b->lds_limit)) return false;
    geo[r] = int2{ty, cz};
    *threads = std::max(*threads, S * (ty + 4));
    *lds = std::max(*lds, mc_slab_lds_bytes(P, ty));
    for (int z0 = 0; z0 < P.nz; z0 += cz)
      for (int y0 = 0; y0 < P.ny; y0 += ty) wg.push_back(int4{r, y0, z0, 0});
  }
  *threads = (*threads + 63) / 64 * 64;
  // Workgroup b runs on XCD b % 8: each XCD takes a contiguous run of the
  // table, so tiles of one replica that run side by side share their halo
  // rows in one L2.
  const int n = (int)wg.size();
  int start[9] = {0};
  for (int x = 0; x < 8; ++x) start[x + 1] = start[x] + (n - x + 7) / 8;  // workgroups b < n with b % 8 == x
  std::vector<int4> perm((size_t)n);
  for (int b = 0; b < n; ++b) perm[(size_t)b] = wg[(size_t)(start[b % 8] + b / 8)];
  wg.swap(perm);
  return true;
}

// Measurement tiles of every replica: ty the largest divisor of Ny whose
// TY + 2 rows fit kMeasureBudget threads, cz the largest divisor of Nz up
// to kMeasureCz; one table entry per tile, each replica's entries
// consecutive.  (On the 64-replica batch, budgets 256-512 and cz 8-32 timed
// within 3 % of each other; 384 was 15 % slower.)
constexpr int kMeasureBudget = 256;
constexpr int kMeasureCz = 16;
bool measure_plan(const mc_batch *b, std::vector<int4> &wg, std::vector<int2> &geo, std::vector<int> &wg0,
                  std::vector<int> &nwg) {
  wg.clear();
  geo.assign(b->nrep, int2{0, 0});
  wg0.assign(b->nrep, 0);
  nwg.assign(b->nrep, 0);
  for (int r = 0; r < b->nrep; ++r) {
    const MCDevParams &P = b->hp[r];
    int ty = 0, cz = 0;
    for (int t = P.ny; t >= 1; --t)
      if (P.ny % t == 0 && (P.nx / 4) * (t + 2) <= kMeasureBudget &&
          mc_measure_slab_lds_bytes(P, t) + 256 <= b->lds_limit) {
        ty = t;
        break;
      }
    for (int c = std::min(kMeasureCz, P.nz); c >= 1; --c)
      if (P.nz % c == 0) {
        cz = c;
        break;
      }
    if (ty < 1 || cz < 1) return false;
    geo[r] = int2{ty, cz};
    wg0[r] = (int)wg.size();
    for (int z0 = 0; z0 < P.nz; z0 += cz)
      for (int y0 = 0; y0 < P.ny; y0 += ty) wg.push_back(int4{r, y0, z0, 0});
    nwg[r] = (int)wg.size() - wg0[r];
  }
  return true;
}

// E, Q and S of every replica into d_out (and series when given): on the
// measurement tiles when they exist, else per segment tile.
void launch_batch_measure(mc_batch *b, double *series) {
  if (b->d_mwg)
    mc_launch_batch_measure_slab(b->launch(), b->M, b->mthreads, b->mlds, b->d_mwg0, b->d_mnwg, b->d_partials,
                                 b->d_out, series);
  else
    mc_launch_batch_measure(b->launch(), b->cu_count * 8, b->d_partials, b->d_out, series);
}

int persist_prepare(mc_batch *b);

// Makes (budget, cz_target, aligned) the slab shape: uploads the tables.
int set_slab_shape(mc_batch *b, int budget, int cz_target, bool aligned) {
  std::vector<int4> wg;
  std::vector<int2> geo;
  int threads = 0;
  size_t lds = 0;
  if (!slab_plan(b, budget, cz_target, aligned, wg, geo, &threads, &lds))
    return bfail(MC_EINVAL, "batched slab: no tile shape fits every replica");
  if (b->d_wg) MB_HIP(hipFree(b->d_wg));
  b->d_wg = nullptr;
  MB_HIP(hipMalloc(&b->d_wg, wg.size() * sizeof(int4)));
  if (!b->d_geo) MB_HIP(hipMalloc(&b->d_geo, b->nrep * sizeof(int2)));
  MB_HIP(hipMemcpy(b->d_wg, wg.data(), wg.size() * sizeof(int4), hipMemcpyHostToDevice));
  MB_HIP(hipMemcpy(b->d_geo, geo.data(), b->nrep * sizeof(int2), hipMemcpyHostToDevice));
  b->S = MCBatchSlab{b->d_wg, b->d_geo, (int)wg.size(), threads, lds};
  b->slab_budget = budget;
  b->slab_cz = cz_target;
  b->slab_aligned = aligned;
  (void)mc_slab_prepare(lds);
  return b->persist ? persist_prepare(b) : MC_OK;
}

// Sweeps per persistent launch (the graphs hold at most this many sweeps).
constexpr int kPersistChunk = 64;

// The persistent kernel's grid (what the device holds at once at the current
// slab shape, at most one workgroup per tile) and its sync words.
int persist_prepare(mc_batch *b) {
  const int resident = mc_batch_persist_grid(b->potential, b->S.threads, b->S.lds, b->cu_count);
  if (resident < 1) return bfail(MC_EINVAL, "persistent batch sweep: the slab tile does not fit a CU");
  b->persist_grid = std::min(resident, b->S.nwg);
  if (!b->d_sync) {
    MB_HIP(hipMalloc(&b->d_sync, (kPersistChunk + 2) * sizeof(unsigned int)));
    MB_HIP(hipMemset(b->d_sync, 0, (kPersistChunk + 2) * sizeof(unsigned int)));
  }
  return MC_OK;
}

// Resolves the requested kernel: the slab when every replica has a tile
// shape and the second plane set can be allocated, else the segment kernel.
// AUTO takes the slab for Lebwohl-Lasher only (GHRL and pear evaluate every
// move exactly, which the segment kernel's occupancy serves better).
int select_kernel(mc_batch *b, int kind) {
  const bool persist = kind == MC_KERNEL_SLAB_PERSISTENT;
  if (persist) kind = MC_KERNEL_SLAB;
  bool want = kind == MC_KERNEL_SLAB || (kind == MC_KERNEL_AUTO && b->potential == MCD_LL);
  if (want && b->slab_budget == 0) {
    std::vector<int4> wg;
    std::vector<int2> geo;
    int threads = 0;
    size_t lds = 0;
    if (slab_plan(b, 256, 16, false, wg, geo, &threads, &lds)) {
      const int rc = set_slab_shape(b, 256, 16, false);
      if (rc) return rc;
    } else {
      want = false;
    }
  }
  if (want && !b->bx) {
    const size_t fb = (size_t)b->total_sites * sizeof(float);
    float *p[3] = {nullptr, nullptr, nullptr};
    for (int d = 0; d < 3; ++d)
      if (hipMalloc(&p[d], fb) != hipSuccess) {
        (void)hipGetLastError();
        for (int e = 0; e < 3; ++e)
          if (p[e]) (void)hipFree(p[e]);
        want = false;
        break;
      }
    if (want) {
      b->bx = p[0];
      b->by = p[1];
      b->bz = p[2];
    }
  }
  if (kind == MC_KERNEL_SLAB && !want)
    return bfail(MC_EINVAL, "mc_batch_set_kernel: no slab tile shape fits every replica (or no memory for the "
                            "second plane set)");
  if (persist) {
    const int rc = persist_prepare(b);
    if (rc) return rc;
  }
  b->kind = persist ? MC_KERNEL_SLAB_PERSISTENT : kind;
  b->slab = want;
  b->persist = persist;
  drop_graphs(b);
  return MC_OK;
}

// n sweeps (n <= kPersistChunk for the persistent kernel).
void enqueue_batch_sweeps(mc_batch *b, int n) {
  if (b->slab && b->persist) {
    mc_launch_batch_slab_persist(b->launch(), b->S, b->bx, b->by, b->bz, b->persist_grid, n, b->d_sync,
                                 b->spin_limit);
    if (n & 1) swap_planes(b);
    return;
  }
  for (int w = 0; w < n; ++w) {
    const MCBatchLaunch B = b->launch();
    if (b->slab) {
      mc_launch_batch_slab(B, b->S, b->bx, b->by, b->bz);
      swap_planes(b);
      continue;
    }
    mc_launch_batch_x4(B, 0, b->sweep_grid(), 0);
    mc_launch_batch_x4(B, 1, b->sweep_grid(), 1);
  }
}

void enqueue_batch_sweep(mc_batch *b) { enqueue_batch_sweeps(b, 1); }

// Reports (and clears) a timed-out grid barrier of a persistent launch; the
// stream must have drained.
int check_fault(mc_batch *b) {
  if (!b->d_sync) return MC_OK;
  unsigned int f = 0;
  MB_HIP(hipMemcpy(&f, b->d_sync + 1, sizeof(f), hipMemcpyDeviceToHost));
  if (!f) return MC_OK;
  const unsigned int zero = 0;
  MB_HIP(hipMemcpy(b->d_sync + 1, &zero, sizeof(zero), hipMemcpyHostToDevice));
  return bfail(MC_EHIP, "persistent batch sweep: a grid barrier did not complete (workgroups not resident at once, "
                        "or the wait limit too small); the lattices are not valid");
}

// Graph of n sweeps from the current plane set.  Capture records launches
// without running them, so the plane swaps made while capturing are undone;
// launch_graph applies them as the replay does.
int get_graph(mc_batch *b, int n, hipGraphExec_t *out) {
  const int key = 2 * n + b->flip;
  auto it = b->graphs.find(key);
  if (it == b->graphs.end()) {
    const int flip0 = b->flip;
    hipGraph_t g = nullptr;
    MB_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    enqueue_batch_sweeps(b, n);
    if (b->flip != flip0) swap_planes(b);
    MB_HIP(hipStreamEndCapture(b->stream, &g));
    hipGraphExec_t ge = nullptr;
    hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    MB_HIP(e);
    it = b->graphs.emplace(key, ge).first;
  }
  *out = it->second;
  return MC_OK;
}

int launch_graph(mc_batch *b, int n, hipGraphExec_t ge) {
  MB_HIP(hipGraphLaunch(ge, b->stream));
  if (b->slab && (n & 1)) swap_planes(b);
  return MC_OK;
}

// n sweeps as graphs of at most kGraphChunk sweeps (the full chunk replayed
// n / kGraphChunk times, then the remainder), so no graph grows with n and
// the cache holds at most 2 * kGraphChunk graphs.  The chunk is even, so the
// remainder graph starts from the plane set it was captured on.
constexpr int kGraphChunk = 64;
static_assert(kGraphChunk % 2 == 0, "the chunk must keep the slab kernel's plane set");
static_assert(kGraphChunk <= kPersistChunk, "a graph's sweeps must fit one persistent launch");

int run_graphs(mc_batch *b, int n, bool prepare_only) {
  const int reps = n / kGraphChunk, rem = n % kGraphChunk;
  hipGraphExec_t chunk = nullptr, rest = nullptr;
  int rc = MC_OK;
  if (reps) rc = get_graph(b, kGraphChunk, &chunk);
  if (rc == MC_OK && rem) rc = get_graph(b, rem, &rest);
  if (rc || prepare_only) return rc;
  for (int r = 0; r < reps && rc == MC_OK; ++r) rc = launch_graph(b, kGraphChunk, chunk);
  if (rc == MC_OK && rem) rc = launch_graph(b, rem, rest);
  return rc;
}

}  // namespace

extern "C" {

int mc_batch_create(int nrep, const mc_params *params, const float *const *ni_aos, mc_batch **out) {
  if (nrep < 1 || !params || !ni_aos || !out) return bfail(MC_EINVAL, "mc_batch_create: bad arguments");
  *out = nullptr;
  for (int r = 0; r < nrep; ++r) {
    const mc_params &p = params[r];
    if (p.nx < 4 || p.ny < 2 || p.nz < 2 || p.nx % 4 || p.ny % 2 || p.nz % 2)
      return bfail(MC_EINVAL, "mc_batch_create: replica " + std::to_string(r) +
                                  " needs Nx % 4 == 0 and even Ny, Nz");
    if (p.bound[0] != MC_PERIODIC || p.bound[1] != MC_PERIODIC || p.bound[2] != MC_PERIODIC)
      return bfail(MC_EINVAL, "mc_batch_create: replicas must be periodic on every axis");
    if (p.n_surf != 0) return bfail(MC_EINVAL, "mc_batch_create: replicas are bulk lattices (no surfaces)");
    if (p.potential != params[0].potential)
      return bfail(MC_EINVAL, "mc_batch_create: every replica must use the same potential");
    if (p.potential < MC_LL || p.potential > MC_PEAR) return bfail(MC_EINVAL, "mc_batch_create: unknown potential");
    if (!(p.T > 0.0f) || !isfinite(p.T)) return bfail(MC_EINVAL, "mc_batch_create: temperatures must be > 0");
    if ((long long)p.nx * p.ny * p.nz >= (1LL << 31)) return bfail(MC_EINVAL, "mc_batch_create: replica too large");
    if (!ni_aos[r]) return bfail(MC_EINVAL, "mc_batch_create: null director array");
    const long long nt = (long long)p.nx * p.ny * p.nz;
    for (long long s = 0; s < nt; ++s) {
      const float *v = ni_aos[r] + 3 * s;
      const float n2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
      if (!isfinite(n2) || fabsf(n2 - 1.0f) > 1e-3f)
        return bfail(MC_EINVAL, "mc_batch_create: replica " + std::to_string(r) + " has a non-unit director");
    }
  }
  if (mc_device_count() < 1) return bfail(MC_ENODEV, "mc_batch_create: no gfx950 device visible");
  mc_batch *b = new mc_batch();
  auto bail = [&](int code) {
    free_batch(b);
    return code;
  };
#define MB_HIPC(call)                                                                          \
  do {                                                                                         \
    hipError_t e_ = (call);                                                                    \
    if (e_ != hipSuccess) {                                                                    \
      (void)hipGetLastError();                                                                 \
      return bail(bfail(e_ == hipErrorOutOfMemory ? MC_ENOMEM : MC_EHIP,                       \
                        std::string(#call) + " failed: " + hipGetErrorString(e_)));           \
    }                                                                                          \
  } while (0)
  MB_HIPC(hipGetDevice(&b->device));
  hipDeviceProp_t prop;
  MB_HIPC(hipGetDeviceProperties(&prop, b->device));
  b->cu_count = prop.multiProcessorCount;
  b->lds_limit = prop.sharedMemPerBlock;
  MB_HIPC(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  MB_HIPC(hipEventCreate(&b->ev0));
  MB_HIPC(hipEventCreate(&b->ev1));
  b->nrep = nrep;
  b->potential = params[0].potential;
  b->hp.resize(nrep);
  b->off.resize(nrep);
  b->tile0.resize(nrep);
  b->ntile.resize(nrep);
  std::vector<int> tile_rep;
  long long maxnt = 0;
  for (int r = 0; r < nrep; ++r) {
    mc_to_dev_params(&params[r], &b->hp[r]);
    const long long nt = (long long)params[r].nx * params[r].ny * params[r].nz;
    b->hp[r].nt = (int)nt;
    b->hp[r].valid = (int)nt;
    b->off[r] = b->total_sites;
    b->total_sites += nt;
    maxnt = std::max(maxnt, nt);
    const int segs = (int)(nt / 4);
    b->tile0[r] = b->ntiles;
    b->ntile[r] = (segs + MCD_BLOCK - 1) / MCD_BLOCK;
    for (int t = 0; t < b->ntile[r]; ++t) tile_rep.push_back(r);
    b->ntiles += b->ntile[r];
  }
  const size_t fb = (size_t)b->total_sites * sizeof(float);
  MB_HIPC(hipMalloc(&b->nx, fb));
  MB_HIPC(hipMalloc(&b->ny, fb));
  MB_HIPC(hipMalloc(&b->nz, fb));
  MB_HIPC(hipMalloc(&b->staging, 3 * (size_t)maxnt * sizeof(float)));
  MB_HIPC(hipMalloc(&b->d_params, nrep * sizeof(MCDevParams)));
  MB_HIPC(hipMalloc(&b->d_states, nrep * sizeof(MCDevState)));
  MB_HIPC(hipMalloc(&b->d_tile_rep, b->ntiles * sizeof(int)));
  MB_HIPC(hipMalloc(&b->d_rep_tile0, nrep * sizeof(int)));
  MB_HIPC(hipMalloc(&b->d_rep_ntiles, nrep * sizeof(int)));
  MB_HIPC(hipMalloc(&b->d_rep_off, nrep * sizeof(long long)));
  MB_HIPC(hipMalloc(&b->d_done, sizeof(uint32_t)));
  {
    std::vector<int4> mwg;
    std::vector<int2> mgeo;
    std::vector<int> mwg0, mnwg;
    const bool mok = measure_plan(b, mwg, mgeo, mwg0, mnwg);
    MB_HIPC(hipMalloc(&b->d_partials, (size_t)std::max(b->ntiles, mok ? (int)mwg.size() : 0) * 8 * sizeof(double)));
    if (mok) {
      MB_HIPC(hipMalloc(&b->d_mwg, mwg.size() * sizeof(int4)));
      MB_HIPC(hipMalloc(&b->d_mgeo, nrep * sizeof(int2)));
      MB_HIPC(hipMalloc(&b->d_mwg0, nrep * sizeof(int)));
      MB_HIPC(hipMalloc(&b->d_mnwg, nrep * sizeof(int)));
      MB_HIPC(hipMemcpy(b->d_mwg, mwg.data(), mwg.size() * sizeof(int4), hipMemcpyHostToDevice));
      MB_HIPC(hipMemcpy(b->d_mgeo, mgeo.data(), nrep * sizeof(int2), hipMemcpyHostToDevice));
      MB_HIPC(hipMemcpy(b->d_mwg0, mwg0.data(), nrep * sizeof(int), hipMemcpyHostToDevice));
      MB_HIPC(hipMemcpy(b->d_mnwg, mnwg.data(), nrep * sizeof(int), hipMemcpyHostToDevice));
      b->mlds = mc_batch_measure_slab_shape(b->hp.data(), mgeo.data(), nrep, &b->mthreads);
      b->M = MCBatchSlab{b->d_mwg, b->d_mgeo, (int)mwg.size(), b->mthreads, b->mlds};
      (void)mc_slab_prepare(b->mlds);
    }
  }
  MB_HIPC(hipMalloc(&b->d_out, (size_t)nrep * 8 * sizeof(double)));
  MB_HIPC(hipMemcpy(b->d_params, b->hp.data(), nrep * sizeof(MCDevParams), hipMemcpyHostToDevice));
  MB_HIPC(hipMemcpy(b->d_tile_rep, tile_rep.data(), b->ntiles * sizeof(int), hipMemcpyHostToDevice));
  MB_HIPC(hipMemcpy(b->d_rep_tile0, b->tile0.data(), nrep * sizeof(int), hipMemcpyHostToDevice));
  MB_HIPC(hipMemcpy(b->d_rep_ntiles, b->ntile.data(), nrep * sizeof(int), hipMemcpyHostToDevice));
  MB_HIPC(hipMemcpy(b->d_rep_off, b->off.data(), nrep * sizeof(long long), hipMemcpyHostToDevice));
  MB_HIPC(hipMemsetAsync(b->d_done, 0, sizeof(uint32_t), b->stream));
  for (int r = 0; r < nrep; ++r) {
    const int nt = b->hp[r].nt;
    MB_HIPC(hipMemcpyAsync(b->staging, ni_aos[r], 3 * (size_t)nt * sizeof(float), hipMemcpyHostToDevice,
                           b->stream));
    mc_launch_aos_to_soa(b->stream, b->staging, b->nx + b->off[r], b->ny + b->off[r], b->nz + b->off[r], nt,
                         std::min(b->cu_count * 8, (nt + 255) / 256));
    MB_HIPC(hipStreamSynchronize(b->stream));
  }
  std::vector<float> T(nrep), E(nrep), A(nrep);
  for (int r = 0; r < nrep; ++r) {
    T[r] = params[r].T;
    E[r] = params[r].elecE;
    A[r] = params[r].ang_var;
  }
  float *tmp = nullptr;
  MB_HIPC(hipMalloc(&tmp, 3 * nrep * sizeof(float)));
  hipError_t e1 = hipMemcpy(tmp, T.data(), nrep * sizeof(float), hipMemcpyHostToDevice);
  if (e1 == hipSuccess) e1 = hipMemcpy(tmp + nrep, E.data(), nrep * sizeof(float), hipMemcpyHostToDevice);
  if (e1 == hipSuccess) e1 = hipMemcpy(tmp + 2 * nrep, A.data(), nrep * sizeof(float), hipMemcpyHostToDevice);
  if (e1 == hipSuccess) {
    k_batch_init<<<(nrep + 255) / 256, 256, 0, b->stream>>>(b->d_states, nrep, tmp, tmp + nrep, tmp + 2 * nrep);
    e1 = hipStreamSynchronize(b->stream);
  }
  (void)hipFree(tmp);
  MB_HIPC(e1);
#undef MB_HIPC
  const int krc = select_kernel(b, MC_KERNEL_AUTO);
  if (krc) return bail(krc);
  *out = b;
  return MC_OK;
}

int mc_batch_destroy(mc_batch *b) {
  if (!b) return MC_OK;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  free_batch(b);
  return MC_OK;
}

int mc_batch_set_T(mc_batch *b, int rep, float T) {
  int rc = check_batch(b);
  if (rc) return rc;
  if (rep < 0 || rep >= b->nrep) return bfail(MC_EINVAL, "mc_batch_set_T: replica out of range");
  if (!(T > 0.0f) || !isfinite(T)) return bfail(MC_EINVAL, "mc_batch_set_T: temperature must be finite and > 0");
  k_batch_set_T<<<1, 1, 0, b->stream>>>(b->d_states, rep, T);
  MB_HIP(hipGetLastError());
  return MC_OK;
}

int mc_batch_sweeps(mc_batch *b, int n) {
  int rc = check_batch(b);
  if (rc) return rc;
  if (n < 0) return bfail(MC_EINVAL, "mc_batch_sweeps: negative sweep count");
  if (n == 0) return MC_OK;
  return run_graphs(b, n, false);
}

int mc_batch_sync(mc_batch *b) {
  int rc = check_batch(b);
  if (rc) return rc;
  MB_HIP(hipStreamSynchronize(b->stream));
  return check_fault(b);
}

int mc_batch_time_sweeps(mc_batch *b, int n, float *ms) {
  int rc = check_batch(b);
  if (rc) return rc;
  if (n < 1 || !ms) return bfail(MC_EINVAL, "mc_batch_time_sweeps: need n >= 1 and an output");
  rc = run_graphs(b, n, true);  // capture and instantiate outside the timed region
  if (rc) return rc;
  MB_HIP(hipEventRecord(b->ev0, b->stream));
  rc = run_graphs(b, n, false);
  if (rc) return rc;
  MB_HIP(hipEventRecord(b->ev1, b->stream));
  MB_HIP(hipEventSynchronize(b->ev1));
  MB_HIP(hipEventElapsedTime(ms, b->ev0, b->ev1));
  return check_fault(b);
}

int mc_batch_measure(mc_batch *b, double *E, double *S) {
  int rc = check_batch(b);
  if (rc) return rc;
  launch_batch_measure(b, nullptr);
  MB_HIP(hipGetLastError());
  std::vector<double> o((size_t)b->nrep * 8);
  MB_HIP(hipMemcpyAsync(o.data(), b->d_out, o.size() * sizeof(double), hipMemcpyDeviceToHost, b->stream));
  MB_HIP(hipStreamSynchronize(b->stream));
  rc = check_fault(b);
  if (rc) return rc;
  for (int r = 0; r < b->nrep; ++r) {
    if (E) E[r] = o[(size_t)r * 8];
    if (S) S[r] = o[(size_t)r * 8 + 1];
  }
  return MC_OK;
}

int mc_batch_measure_series(mc_batch *b, int n, double *E, double *S) {
  int rc = check_batch(b);
  if (rc) return rc;
  if (n < 1) return bfail(MC_EINVAL, "mc_batch_measure_series: need n >= 1");
  const size_t need = 2 * (size_t)n * b->nrep;
  if ((int)need > b->series_cap) {
    if (b->d_series) MB_HIP(hipFree(b->d_series));
    b->d_series = nullptr;
    b->series_cap = 0;
    MB_HIP(hipMalloc(&b->d_series, need * sizeof(double)));
    b->series_cap = (int)need;
  }
  for (int w = 0; w < n; ++w) {
    enqueue_batch_sweep(b);
    // (after the sweep: the slab kernel moves the lattice to the other plane set)
    launch_batch_measure(b, b->d_series + 2 * (size_t)w * b->nrep);
  }
  MB_HIP(hipGetLastError());
  std::vector<double> o(need);
  MB_HIP(hipMemcpyAsync(o.data(), b->d_series, need * sizeof(double), hipMemcpyDeviceToHost, b->stream));
  MB_HIP(hipStreamSynchronize(b->stream));
  rc = check_fault(b);
  if (rc) return rc;
  for (int w = 0; w < n; ++w)
    for (int r = 0; r < b->nrep; ++r) {
      if (E) E[(size_t)w * b->nrep + r] = o[2 * ((size_t)w * b->nrep + r)];
      if (S) S[(size_t)w * b->nrep + r] = o[2 * ((size_t)w * b->nrep + r) + 1];
    }
  return MC_OK;
}

int mc_batch_get_directors_aos(mc_batch *b, int rep, float *ni_aos) {
  int rc = check_batch(b);
  if (rc) return rc;
  if (rep < 0 || rep >= b->nrep || !ni_aos) return bfail(MC_EINVAL, "mc_batch_get_directors_aos: bad arguments");
  const int nt = b->hp[rep].nt;
  mc_launch_soa_to_aos(b->stream, b->nx + b->off[rep], b->ny + b->off[rep], b->nz + b->off[rep], b->staging, nt,
                       std::min(b->cu_count * 8, (nt + 255) / 256));
  MB_HIP(hipGetLastError());
  MB_HIP(hipMemcpyAsync(ni_aos, b->staging, 3 * (size_t)nt * sizeof(float), hipMemcpyDeviceToHost, b->stream));
  MB_HIP(hipStreamSynchronize(b->stream));
  return MC_OK;
}

int mc_batch_get_state(mc_batch *b, uint32_t *accepted, uint32_t *sweeps, float *ang_var) {
  int rc = check_batch(b);
  if (rc) return rc;
  std::vector<MCDevState> st(b->nrep);
  MB_HIP(hipMemcpyAsync(st.data(), b->d_states, b->nrep * sizeof(MCDevState), hipMemcpyDeviceToHost, b->stream));
  MB_HIP(hipStreamSynchronize(b->stream));
  rc = check_fault(b);
  if (rc) return rc;
  for (int r = 0; r < b->nrep; ++r) {
    if (accepted) accepted[r] = st[r].last_acc;
    if (sweeps) sweeps[r] = st[r].sweep;
    if (ang_var) ang_var[r] = st[r].ang_var;
  }
  return MC_OK;
}

int mc_batch_set_tuning(mc_batch *b, int grid, int xcd) {
  int rc = check_batch(b);
  if (rc) return rc;
  if (grid < 0 || (xcd != 0 && xcd != 1) || (xcd && grid % 8))
    return bfail(MC_EINVAL, "mc_batch_set_tuning: grid must be >= 0 and a multiple of 8 with xcd = 1");
  MB_HIP(hipStreamSynchronize(b->stream));
  drop_graphs(b);
  b->grid = grid;
  b->xcd = xcd;
  b->kind = MC_KERNEL_X4;
  b->slab = false;
  b->persist = false;
  return MC_OK;
}

int mc_batch_set_kernel(mc_batch *b, int kind) {
  int rc = check_batch(b);
  if (rc) return rc;
  if (kind != MC_KERNEL_AUTO && kind != MC_KERNEL_SLAB && kind != MC_KERNEL_X4 && kind != MC_KERNEL_SLAB_PERSISTENT)
    return bfail(MC_EINVAL, "mc_batch_set_kernel: kind must be MC_KERNEL_AUTO, MC_KERNEL_SLAB, "
                            "MC_KERNEL_SLAB_PERSISTENT or MC_KERNEL_X4");
  MB_HIP(hipStreamSynchronize(b->stream));
  return select_kernel(b, kind);
}

int mc_batch_get_kernel(mc_batch *b, int *kind) {
  int rc = check_batch(b);
  if (rc) return rc;
  if (!kind) return bfail(MC_EINVAL, "mc_batch_get_kernel: null output");
  *kind = b->slab ? (b->persist ? MC_KERNEL_SLAB_PERSISTENT : MC_KERNEL_SLAB) : MC_KERNEL_X4;
  return MC_OK;
}

int mc_batch_set_barrier_limit(mc_batch *b, unsigned int polls) {
  int rc = check_batch(b);
  if (rc) return rc;
  if (polls < 1) return bfail(MC_EINVAL, "mc_batch_set_barrier_limit: need at least one poll");
  MB_HIP(hipStreamSynchronize(b->stream));
  drop_graphs(b);
  b->spin_limit = polls;
  return MC_OK;
}

int mc_batch_autotune(mc_batch *b, int *grid, int *xcd, float *best_us) {
  // Candidates sweep scratch copies of the planes and states, so tuning
  // leaves every chain untouched: segment-kernel grids, then (when the batch
  // may use it) slab shapes over thread budgets and planes per tile.
  int rc = check_batch(b);
  if (rc) return rc;
  const size_t nt = (size_t)b->total_sites;
  const bool try_slab = b->kind != MC_KERNEL_X4 && b->bx != nullptr;
  float *scratch = nullptr;
  MCDevState *sst = nullptr;
  int4 *twg = nullptr;
  int2 *tgeo = nullptr;
  MB_HIP(hipMalloc(&scratch, (try_slab ? 6 : 3) * nt * sizeof(float)));
  auto release = [&]() {
    (void)hipFree(scratch);
    if (sst) (void)hipFree(sst);
    if (twg) (void)hipFree(twg);
    if (tgeo) (void)hipFree(tgeo);
  };
  if (hipMalloc(&sst, b->nrep * sizeof(MCDevState)) != hipSuccess ||
      hipMalloc(&tgeo, b->nrep * sizeof(int2)) != hipSuccess) {
    (void)hipGetLastError();
    release();
    return bfail(MC_ENOMEM, "mc_batch_autotune: scratch allocation failed");
  }
  MCBatchLaunch B = b->launch();
  B.states = sst;
  hipError_t err = hipSuccess;
  auto reset = [&]() {
    const float *src[3] = {b->nx, b->ny, b->nz};
    for (int d = 0; d < 3 && err == hipSuccess; ++d)
      err = hipMemcpyAsync(scratch + d * nt, src[d], nt * sizeof(float), hipMemcpyDeviceToDevice, b->stream);
    if (err == hipSuccess)
      err = hipMemcpyAsync(sst, b->d_states, b->nrep * sizeof(MCDevState), hipMemcpyDeviceToDevice, b->stream);
  };
  auto time_it = [&](auto &&sweep) {
    for (int w = 0; w < 2; ++w) sweep();
    const int reps = 6;
    float ms = 0.0f;
    if (err == hipSuccess) err = hipEventRecord(b->ev0, b->stream);
    for (int w = 0; w < reps; ++w) sweep();
    if (err == hipSuccess) err = hipGetLastError();
    if (err == hipSuccess) err = hipEventRecord(b->ev1, b->stream);
    if (err == hipSuccess) err = hipEventSynchronize(b->ev1);
    if (err == hipSuccess) err = hipEventElapsedTime(&ms, b->ev0, b->ev1);
    return ms * 1e3f / reps;
  };
  int best_g = 0, best_x = 0, best_budget = 0, best_cz = 0;
  bool best_al = false;
  float best = 1e30f;
  const int mults[] = {2, 4, 8, 16, 32};
  for (int xm = 0; xm < 2 && err == hipSuccess; ++xm)
    for (int m : mults) {
      const int g = b->cu_count * m;
      if (xm && g % 8) continue;
      B.xcd = xm;
      B.nx = scratch;
      B.ny = scratch + nt;
      B.nz = scratch + 2 * nt;
      reset();
      if (err != hipSuccess) break;
      const float us = time_it([&]() {
        mc_launch_batch_x4(B, 0, g, 0);
        mc_launch_batch_x4(B, 1, g, 1);
      });
      if (err != hipSuccess) break;
      if (us < best) {
        best = us;
        best_g = g;
        best_x = xm;
      }
    }
  if (try_slab) {
    const int budgets[] = {128, 192, 256, 384, 512, 768};
    const int czs[] = {8, 16, 32};
    for (int budget : budgets)
      for (int czt : czs)
        for (int al = 0; al < 2; ++al) {
          if (err != hipSuccess) break;
          std::vector<int4> wg;
          std::vector<int2> geo;
          int threads = 0;
          size_t lds = 0;
          if (!slab_plan(b, budget, czt, al != 0, wg, geo, &threads, &lds)) continue;
          if (twg) (void)hipFree(twg);
          twg = nullptr;
          err = hipMalloc(&twg, wg.size() * sizeof(int4));
          if (err == hipSuccess)
            err = hipMemcpy(twg, wg.data(), wg.size() * sizeof(int4), hipMemcpyHostToDevice);
          if (err == hipSuccess) err = hipMemcpy(tgeo, geo.data(), b->nrep * sizeof(int2), hipMemcpyHostToDevice);
          if (err != hipSuccess) break;
          (void)mc_slab_prepare(lds);
          const MCBatchSlab S{twg, tgeo, (int)wg.size(), threads, lds};
          reset();
          if (err != hipSuccess) break;
          int cur = 0;
          B.xcd = 0;
          const float us = time_it([&]() {
            float *io = scratch + (size_t)cur * 3 * nt, *oo = scratch + (size_t)(cur ^ 1) * 3 * nt;
            B.nx = io;
            B.ny = io + nt;
            B.nz = io + 2 * nt;
            mc_launch_batch_slab(B, S, oo, oo + nt, oo + 2 * nt);
            cur ^= 1;
          });
          if (err != hipSuccess) break;
          if (us < best) {
            best = us;
            best_budget = budget;
            best_cz = czt;
            best_al = al != 0;
            best_g = (int)wg.size();
            best_x = 0;
          }
        }
  }
  release();
  MB_HIP(err);
  if (best_budget > 0) {
    MB_HIP(hipStreamSynchronize(b->stream));
    rc = set_slab_shape(b, best_budget, best_cz, best_al);
    if (rc) return rc;
    b->slab = true;
    drop_graphs(b);
  } else {
    rc = mc_batch_set_tuning(b, best_g, best_x);
    if (rc) return rc;
  }
  if (grid) *grid = best_g;
  if (xcd) *xcd = best_x;
  if (best_us) *best_us = best;
  return MC_OK;
}

}  // extern "C"
