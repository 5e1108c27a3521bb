// C ABI of the learned proposal (declared in include/mclics_hip.h): the
// network's parameters, training data drawn from plain-sampler lattices,
// bf16-MFMA training with Adam, inference, and the learned
// Metropolis-Hastings sweeps of a lattice handle.  No reference counterpart.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/mclics_hip.h"
#include "mc_mlp.h"

namespace {

int lfail(int code, const std::string &msg) {
  mc_set_error_message(msg.c_str());
  return code;
}

#define ML_HIP(call)                                                                          \
  do {                                                                                        \
    hipError_t e_ = (call);                                                                   \
    if (e_ != hipSuccess) {                                                                   \
      (void)hipGetLastError();                                                                \
      return lfail(e_ == hipErrorOutOfMemory ? MC_ENOMEM : MC_EHIP,                           \
                   std::string(#call) + " failed: " + hipGetErrorString(e_));                \
    }                                                                                         \
  } while (0)

// Initial parameters: weights uniform in +-sqrt(6/(fan_in+fan_out)) from a
// 64-bit LCG-seeded splitmix stream, biases zero, the log-kappa bias 1 and the
// output weights scaled by 0.1 so the first proposals centre on the
// neighbours' mean axis.
void init_params(uint64_t seed, std::vector<float> &P) {
  P.assign(MC_MLP_NPARAM, 0.0f);
  uint64_t x = seed * 0x9E3779B97F4A7C15ull + 0xD1B54A32D192ED03ull;
  auto next = [&]() {  // splitmix64 (Steele, Lea, Flood 2014)
    x += 0x9E3779B97F4A7C15ull;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
  };
  auto fill = [&](int off, int n_out, int n_in, double scale) {
    const double a = scale * sqrt(6.0 / (n_in + n_out));
    for (int t = 0; t < n_out * n_in; ++t) P[off + t] = (float)((2.0 * next() - 1.0) * a);
  };
  fill(0, 64, 22, 1.0);
  fill(64 * 22 + 64, 64, 64, 1.0);
  fill(64 * 22 + 64 + 64 * 64 + 64, 4, 64, 0.1);
  P[MC_MLP_NPARAM - 1] = 1.0f;
}

}  // namespace

struct mc_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  MCMLPBuffers B{};
  float *X = nullptr, *base = nullptr, *dirs = nullptr;  // training set
  int n = 0, cap = 0;
  int step = 0;
  float *tmp = nullptr;  // host-transfer scratch for infer/grad
  size_t tmp_cap = 0;
};

namespace {

void free_mlp(mc_mlp *m) {
  if (!m) return;
  void *ptrs[] = {m->B.params, m->B.grad, m->B.m1, m->B.m2, m->B.loss, m->B.W1b, m->B.W2b, m->B.W3b, m->B.W3Tb,
                  m->B.W2Tb, m->B.b1, m->B.b2, m->B.b3, m->B.ops, m->X, m->base, m->dirs, m->tmp};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}

int check_mlp(mc_mlp *m) {
  if (!m) return lfail(MC_EINVAL, "null network");
  if (hipSetDevice(m->device) != hipSuccess) return lfail(MC_EHIP, "hipSetDevice failed for the network's device");
  return MC_OK;
}

int ensure_tmp(mc_mlp *m, size_t floats) {
  if (floats <= m->tmp_cap) return MC_OK;
  if (m->tmp) ML_HIP(hipFree(m->tmp));
  m->tmp = nullptr;
  m->tmp_cap = 0;
  ML_HIP(hipMalloc(&m->tmp, floats * sizeof(float)));
  m->tmp_cap = floats;
  return MC_OK;
}

int ensure_dataset(mc_mlp *m, int need) {
  if (need <= m->cap) return MC_OK;
  int cap = std::max(need, 2 * m->cap);
  float *X = nullptr, *b = nullptr, *d = nullptr;
  ML_HIP(hipMalloc(&X, (size_t)cap * 22 * sizeof(float)));
  if (hipMalloc(&b, (size_t)cap * 3 * sizeof(float)) != hipSuccess ||
      hipMalloc(&d, (size_t)cap * 3 * sizeof(float)) != hipSuccess) {
    (void)hipFree(X);
    if (b) (void)hipFree(b);
    return lfail(MC_ENOMEM, "mc_mlp: training-set allocation failed");
  }
  if (m->n) {
    ML_HIP(hipMemcpyAsync(X, m->X, (size_t)m->n * 22 * sizeof(float), hipMemcpyDeviceToDevice, m->stream));
    ML_HIP(hipMemcpyAsync(b, m->base, (size_t)m->n * 3 * sizeof(float), hipMemcpyDeviceToDevice, m->stream));
    ML_HIP(hipMemcpyAsync(d, m->dirs, (size_t)m->n * 3 * sizeof(float), hipMemcpyDeviceToDevice, m->stream));
    ML_HIP(hipStreamSynchronize(m->stream));
  }
  if (m->X) (void)hipFree(m->X);
  if (m->base) (void)hipFree(m->base);
  if (m->dirs) (void)hipFree(m->dirs);
  m->X = X;
  m->base = b;
  m->dirs = d;
  m->cap = cap;
  return MC_OK;
}

}  // namespace

extern "C" {

int mc_mlp_create(const float *params, uint64_t seed, mc_mlp **out) {
  if (!out) return lfail(MC_EINVAL, "mc_mlp_create: null output");
  *out = nullptr;
  if (mc_device_count() < 1) return lfail(MC_ENODEV, "mc_mlp_create: no gfx950 device visible");
  mc_mlp *m = new mc_mlp();
  auto bail = [&](int code) {
    free_mlp(m);
    return code;
  };
#define ML_HIPC(call)                                                                            \
  do {                                                                                           \
    hipError_t e_ = (call);                                                                      \
    if (e_ != hipSuccess) {                                                                      \
      (void)hipGetLastError();                                                                   \
      return bail(lfail(e_ == hipErrorOutOfMemory ? MC_ENOMEM : MC_EHIP,                         \
                        std::string(#call) + " failed: " + hipGetErrorString(e_)));             \
    }                                                                                            \
  } while (0)
  ML_HIPC(hipGetDevice(&m->device));
  ML_HIPC(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
  const size_t pb = MC_MLP_NPARAM * sizeof(float);
  ML_HIPC(hipMalloc(&m->B.params, pb));
  ML_HIPC(hipMalloc(&m->B.grad, pb));
  ML_HIPC(hipMalloc(&m->B.m1, pb));
  ML_HIPC(hipMalloc(&m->B.m2, pb));
  ML_HIPC(hipMalloc(&m->B.loss, sizeof(double)));
  ML_HIPC(hipMalloc(&m->B.W1b, 64 * 32 * sizeof(__bf16)));
  ML_HIPC(hipMalloc(&m->B.W2b, 64 * 64 * sizeof(__bf16)));
  ML_HIPC(hipMalloc(&m->B.W3b, 16 * 64 * sizeof(__bf16)));
  ML_HIPC(hipMalloc(&m->B.W3Tb, 64 * 32 * sizeof(__bf16)));
  ML_HIPC(hipMalloc(&m->B.W2Tb, 64 * 64 * sizeof(__bf16)));
  ML_HIPC(hipMalloc(&m->B.b1, 64 * sizeof(float)));
  ML_HIPC(hipMalloc(&m->B.b2, 64 * sizeof(float)));
  ML_HIPC(hipMalloc(&m->B.b3, 4 * sizeof(float)));
  ML_HIPC(hipMalloc(&m->B.ops, MC_MLP_OPS_BYTES));
  std::vector<float> P;
  if (params)
    P.assign(params, params + MC_MLP_NPARAM);
  else
    init_params(seed, P);
  ML_HIPC(hipMemcpy(m->B.params, P.data(), pb, hipMemcpyHostToDevice));
  ML_HIPC(hipMemset(m->B.grad, 0, pb));
  ML_HIPC(hipMemset(m->B.m1, 0, pb));
  ML_HIPC(hipMemset(m->B.m2, 0, pb));
  mc_mlp_launch_pack(m->B, m->stream);
  ML_HIPC(hipGetLastError());
  ML_HIPC(hipStreamSynchronize(m->stream));
#undef ML_HIPC
  *out = m;
  return MC_OK;
}

int mc_mlp_destroy(mc_mlp *m) {
  if (!m) return MC_OK;
  (void)hipSetDevice(m->device);
  (void)hipStreamSynchronize(m->stream);
  free_mlp(m);
  return MC_OK;
}

int mc_mlp_get_params(mc_mlp *m, float *params) {
  int rc = check_mlp(m);
  if (rc) return rc;
  if (!params) return lfail(MC_EINVAL, "mc_mlp_get_params: null output");
  ML_HIP(hipMemcpyAsync(params, m->B.params, MC_MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  ML_HIP(hipStreamSynchronize(m->stream));
  return MC_OK;
}

int mc_mlp_set_params(mc_mlp *m, const float *params) {
  int rc = check_mlp(m);
  if (rc) return rc;
  if (!params) return lfail(MC_EINVAL, "mc_mlp_set_params: null input");
  ML_HIP(hipMemcpyAsync(m->B.params, params, MC_MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice, m->stream));
  mc_mlp_launch_pack(m->B, m->stream);
  ML_HIP(hipGetLastError());
  ML_HIP(hipStreamSynchronize(m->stream));
  m->step = 0;
  return MC_OK;
}

int mc_mlp_infer(mc_mlp *m, int n, const float *X, const float *base, float *mu, float *logk) {
  int rc = check_mlp(m);
  if (rc) return rc;
  if (n < 1 || !X || !base || !mu || !logk) return lfail(MC_EINVAL, "mc_mlp_infer: bad arguments");
  rc = ensure_tmp(m, (size_t)n * (22 + 3 + 3 + 1));
  if (rc) return rc;
  float *dX = m->tmp, *db = dX + (size_t)n * 22, *dmu = db + (size_t)n * 3, *dlk = dmu + (size_t)n * 3;
  ML_HIP(hipMemcpyAsync(dX, X, (size_t)n * 22 * sizeof(float), hipMemcpyHostToDevice, m->stream));
  ML_HIP(hipMemcpyAsync(db, base, (size_t)n * 3 * sizeof(float), hipMemcpyHostToDevice, m->stream));
  mc_mlp_launch_infer(m->B, m->stream, dX, db, n, dmu, dlk);
  ML_HIP(hipGetLastError());
  ML_HIP(hipMemcpyAsync(mu, dmu, (size_t)n * 3 * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  ML_HIP(hipMemcpyAsync(logk, dlk, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  ML_HIP(hipStreamSynchronize(m->stream));
  return MC_OK;
}

int mc_mlp_grad(mc_mlp *m, int n, const float *X, const float *base, const float *dirs, double *loss, float *grad) {
  int rc = check_mlp(m);
  if (rc) return rc;
  if (n < 1 || !X || !base || !dirs) return lfail(MC_EINVAL, "mc_mlp_grad: bad arguments");
  rc = ensure_tmp(m, (size_t)n * 28);
  if (rc) return rc;
  float *dX = m->tmp, *db = dX + (size_t)n * 22, *dd = db + (size_t)n * 3;
  ML_HIP(hipMemcpyAsync(dX, X, (size_t)n * 22 * sizeof(float), hipMemcpyHostToDevice, m->stream));
  ML_HIP(hipMemcpyAsync(db, base, (size_t)n * 3 * sizeof(float), hipMemcpyHostToDevice, m->stream));
  ML_HIP(hipMemcpyAsync(dd, dirs, (size_t)n * 3 * sizeof(float), hipMemcpyHostToDevice, m->stream));
  mc_mlp_launch_fill(m->B.grad, MC_MLP_NPARAM, 0.0f, m->stream);
  ML_HIP(hipMemsetAsync(m->B.loss, 0, sizeof(double), m->stream));
  mc_mlp_launch_grad(m->B, m->stream, dX, db, dd, n, std::min(1024, (n + 31) / 32));
  ML_HIP(hipGetLastError());
  std::vector<float> g(MC_MLP_NPARAM);
  double l = 0;
  ML_HIP(hipMemcpyAsync(g.data(), m->B.grad, g.size() * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  ML_HIP(hipMemcpyAsync(&l, m->B.loss, sizeof(double), hipMemcpyDeviceToHost, m->stream));
  mc_mlp_launch_fill(m->B.grad, MC_MLP_NPARAM, 0.0f, m->stream);
  ML_HIP(hipStreamSynchronize(m->stream));
  if (grad)
    for (int p = 0; p < MC_MLP_NPARAM; ++p) grad[p] = g[p] / n;
  if (loss) *loss = l / n;
  return MC_OK;
}

int mc_mlp_dataset_add(mc_mlp *m, mc_handle *h, int n, uint32_t seed) {
  int rc = check_mlp(m);
  if (rc) return rc;
  if (!h || n < 1) return lfail(MC_EINVAL, "mc_mlp_dataset_add: bad arguments");
  MCLaunch L;
  rc = mc_handle_launch(h, &L);
  if (rc) return rc;
  if (L.pt) return lfail(MC_EINVAL, "mc_mlp_dataset_add: training lattices must be all active (bulk)");
  rc = ensure_dataset(m, m->n + n);
  if (rc) return rc;
  ML_HIP(hipStreamSynchronize(m->stream));
  mc_mlp_launch_dataset(L, n, seed, m->X + (size_t)m->n * 22, m->base + (size_t)m->n * 3,
                        m->dirs + (size_t)m->n * 3);
  ML_HIP(hipGetLastError());
  ML_HIP(hipStreamSynchronize(L.stream));
  m->n += n;
  return MC_OK;
}

int mc_mlp_dataset_size(mc_mlp *m, int *n) {
  int rc = check_mlp(m);
  if (rc) return rc;
  if (!n) return lfail(MC_EINVAL, "mc_mlp_dataset_size: null output");
  *n = m->n;
  return MC_OK;
}

int mc_mlp_dataset_clear(mc_mlp *m) {
  int rc = check_mlp(m);
  if (rc) return rc;
  m->n = 0;
  return MC_OK;
}

int mc_mlp_train(mc_mlp *m, int epochs, int batch, float lr, float *loss_per_epoch) {
  int rc = check_mlp(m);
  if (rc) return rc;
  if (epochs < 1 || batch < 32 || batch % 32 || !(lr > 0.0f))
    return lfail(MC_EINVAL, "mc_mlp_train: need epochs >= 1, batch a positive multiple of 32, lr > 0");
  if (m->n < batch) return lfail(MC_ESTATE, "mc_mlp_train: the training set holds fewer samples than one batch");
  const int nb = m->n / batch;
  for (int e = 0; e < epochs; ++e) {
    ML_HIP(hipMemsetAsync(m->B.loss, 0, sizeof(double), m->stream));
    for (int b = 0; b < nb; ++b) {
      const size_t o = (size_t)((b + (size_t)e * 7919) % nb) * batch;  // epoch-rotated batch order
      mc_mlp_launch_grad(m->B, m->stream, m->X + o * 22, m->base + o * 3, m->dirs + o * 3, batch,
                         std::min(1024, batch / 32));
      mc_mlp_launch_adam(m->B, m->stream, batch, lr, ++m->step);
      mc_mlp_launch_pack(m->B, m->stream);
    }
    ML_HIP(hipGetLastError());
    double l = 0;
    ML_HIP(hipMemcpyAsync(&l, m->B.loss, sizeof(double), hipMemcpyDeviceToHost, m->stream));
    ML_HIP(hipStreamSynchronize(m->stream));
    if (loss_per_epoch) loss_per_epoch[e] = (float)(l / ((double)nb * batch));
  }
  return MC_OK;
}

}  // extern "C"

const MCMLPBuffers &mc_mlp_buffers(const mc_mlp *m) { return m->B; }
int mc_mlp_device(const mc_mlp *m) { return m->device; }
