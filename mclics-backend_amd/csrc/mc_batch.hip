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
                  b->ny, b->nz, b->d_done, b->d_partials, b->d_out, b->d_series, b->staging};
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

void enqueue_batch_sweep(mc_batch *b) {
  const MCBatchLaunch B = b->launch();
  mc_launch_batch_x4(B, 0, b->sweep_grid(), 0);
  mc_launch_batch_x4(B, 1, b->sweep_grid(), 1);
}

int get_graph(mc_batch *b, int n, hipGraphExec_t *out) {
  auto it = b->graphs.find(n);
  if (it == b->graphs.end()) {
    hipGraph_t g = nullptr;
    MB_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    for (int w = 0; w < n; ++w) enqueue_batch_sweep(b);
    MB_HIP(hipStreamEndCapture(b->stream, &g));
    hipGraphExec_t ge = nullptr;
    hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    MB_HIP(e);
    it = b->graphs.emplace(n, ge).first;
  }
  *out = it->second;
  return MC_OK;
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
  MB_HIPC(hipMalloc(&b->d_partials, (size_t)b->ntiles * 8 * sizeof(double)));
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
  hipGraphExec_t ge = nullptr;
  rc = get_graph(b, n, &ge);
  if (rc) return rc;
  MB_HIP(hipGraphLaunch(ge, b->stream));
  return MC_OK;
}

int mc_batch_sync(mc_batch *b) {
  int rc = check_batch(b);
  if (rc) return rc;
  MB_HIP(hipStreamSynchronize(b->stream));
  return MC_OK;
}

int mc_batch_time_sweeps(mc_batch *b, int n, float *ms) {
  int rc = check_batch(b);
  if (rc) return rc;
  if (n < 1 || !ms) return bfail(MC_EINVAL, "mc_batch_time_sweeps: need n >= 1 and an output");
  hipGraphExec_t ge = nullptr;
  rc = get_graph(b, n, &ge);
  if (rc) return rc;
  MB_HIP(hipEventRecord(b->ev0, b->stream));
  MB_HIP(hipGraphLaunch(ge, b->stream));
  MB_HIP(hipEventRecord(b->ev1, b->stream));
  MB_HIP(hipEventSynchronize(b->ev1));
  MB_HIP(hipEventElapsedTime(ms, b->ev0, b->ev1));
  return MC_OK;
}

int mc_batch_measure(mc_batch *b, double *E, double *S) {
  int rc = check_batch(b);
  if (rc) return rc;
  mc_launch_batch_measure(b->launch(), b->cu_count * 8, b->d_partials, b->d_out, nullptr);
  MB_HIP(hipGetLastError());
  std::vector<double> o((size_t)b->nrep * 8);
  MB_HIP(hipMemcpyAsync(o.data(), b->d_out, o.size() * sizeof(double), hipMemcpyDeviceToHost, b->stream));
  MB_HIP(hipStreamSynchronize(b->stream));
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
  const MCBatchLaunch B = b->launch();
  for (int w = 0; w < n; ++w) {
    enqueue_batch_sweep(b);
    mc_launch_batch_measure(B, b->cu_count * 8, b->d_partials, b->d_out, b->d_series + 2 * (size_t)w * b->nrep);
  }
  MB_HIP(hipGetLastError());
  std::vector<double> o(need);
  MB_HIP(hipMemcpyAsync(o.data(), b->d_series, need * sizeof(double), hipMemcpyDeviceToHost, b->stream));
  MB_HIP(hipStreamSynchronize(b->stream));
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
  for (auto &kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
  b->graphs.clear();
  b->grid = grid;
  b->xcd = xcd;
  return MC_OK;
}

int mc_batch_autotune(mc_batch *b, int *grid, int *xcd, float *best_us) {
  int rc = check_batch(b);
  if (rc) return rc;
  const size_t nt = (size_t)b->total_sites;
  float *scratch = nullptr;
  MCDevState *sst = nullptr;
  MB_HIP(hipMalloc(&scratch, 3 * nt * sizeof(float)));
  if (hipMalloc(&sst, b->nrep * sizeof(MCDevState)) != hipSuccess) {
    (void)hipFree(scratch);
    return bfail(MC_ENOMEM, "mc_batch_autotune: scratch allocation failed");
  }
  MCBatchLaunch B = b->launch();
  B.nx = scratch;
  B.ny = scratch + nt;
  B.nz = scratch + 2 * nt;
  B.states = sst;
  int best_g = 0, best_x = 0;
  float best = 1e30f;
  const int mults[] = {2, 4, 8, 16, 32};
  hipError_t err = hipSuccess;
  for (int xm = 0; xm < 2 && err == hipSuccess; ++xm)
    for (int m : mults) {
      const int g = b->cu_count * m;
      if (xm && g % 8) continue;
      B.xcd = xm;
      err = hipMemcpyAsync(scratch, b->nx, nt * sizeof(float), hipMemcpyDeviceToDevice, b->stream);
      if (err == hipSuccess) err = hipMemcpyAsync(scratch + nt, b->ny, nt * sizeof(float), hipMemcpyDeviceToDevice, b->stream);
      if (err == hipSuccess) err = hipMemcpyAsync(scratch + 2 * nt, b->nz, nt * sizeof(float), hipMemcpyDeviceToDevice, b->stream);
      if (err == hipSuccess)
        err = hipMemcpyAsync(sst, b->d_states, b->nrep * sizeof(MCDevState), hipMemcpyDeviceToDevice, b->stream);
      if (err != hipSuccess) break;
      for (int w = 0; w < 2; ++w) {
        mc_launch_batch_x4(B, 0, g, 0);
        mc_launch_batch_x4(B, 1, g, 1);
      }
      const int reps = 6;
      err = hipEventRecord(b->ev0, b->stream);
      for (int w = 0; w < reps; ++w) {
        mc_launch_batch_x4(B, 0, g, 0);
        mc_launch_batch_x4(B, 1, g, 1);
      }
      if (err == hipSuccess) err = hipEventRecord(b->ev1, b->stream);
      if (err == hipSuccess) err = hipEventSynchronize(b->ev1);
      float ms = 0.0f;
      if (err == hipSuccess) err = hipEventElapsedTime(&ms, b->ev0, b->ev1);
      if (err != hipSuccess) break;
      const float us = ms * 1e3f / reps;
      if (us < best) {
        best = us;
        best_g = g;
        best_x = xm;
      }
    }
  (void)hipFree(scratch);
  (void)hipFree(sst);
  MB_HIP(err);
  rc = mc_batch_set_tuning(b, best_g, best_x);
  if (rc) return rc;
  if (grid) *grid = best_g;
  if (xcd) *xcd = best_x;
  if (best_us) *best_us = best;
  return MC_OK;
}

}  // extern "C"
