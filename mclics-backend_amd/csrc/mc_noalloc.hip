// mc_measure_noalloc: the measurement of mc_measure (mean site energy, FP64
// Q tensor and S) on a lattice the caller holds in device memory, with no
// device allocation.  The kernels are mc_measure's own (the slab-layout
// measurement, else the segment measurement, then the one-block finisher),
// so for the same lattice the result is mc_measure's bit for bit.
//
// Scratch.  Those kernels need a device state (the field strength they read)
// and one row of eight FP64 partial sums per workgroup.  Both live in the
// file-scope device array `scratch` below: 256 bytes of state, then
// kMaxParts rows of partials (256 KiB in all).  The runtime gives every GPU
// its own copy of the array when the code object is loaded there, and each
// copy stays until the process ends.  In host code `scratch` names a host
// stand-in, not device memory, so its device address is looked up once per
// device (hipGetSymbolAddress, with that device current) and only that
// address is ever handed to a kernel.
//
// Every call on one GPU shares that GPU's copy: calls that overlap on one GPU
// (two streams, or two host threads) race on the partials and must be
// ordered by the caller.  Nothing in the array has to start at zero: each
// workgroup writes its whole row of partials and the finisher reads only the
// rows of the current launch; the state is written by the call itself.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <mutex>
#include <string>

#include "../../include/mclics_hip.h"
#include "mc_kernels.h"

namespace {

constexpr int kMaxParts = MC_NOALLOC_MAX_PARTIALS;
constexpr int kStateDoubles = 32;  // 256 bytes for the MCDevState
static_assert(sizeof(MCDevState) <= kStateDoubles * sizeof(double), "state slot too small");

static __device__ __align__(256) double scratch[kStateDoubles + (size_t)kMaxParts * 8];

constexpr int kMaxDevices = 64;

struct DeviceSlot {
  bool ready = false;
  double *ptr = nullptr;  // device address of this device's `scratch`
  int cu_count = 0;
  size_t lds_limit = 65536;
};

std::mutex g_mutex;
DeviceSlot g_slots[kMaxDevices];

int fail(int code, const std::string &msg) {
  mc_set_error_message(msg.c_str());
  return code;
}

// Looks up (once per device) the address of this device's scratch copy and
// the launch limits the shape choice needs.
int device_slot(DeviceSlot *out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    (void)hipGetLastError();
    return fail(MC_ENODEV, "mc_measure_noalloc: no current HIP device");
  }
  if (dev < 0 || dev >= kMaxDevices) return fail(MC_ENODEV, "mc_measure_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_mutex);
  DeviceSlot &s = g_slots[dev];
  if (!s.ready) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) {
      (void)hipGetLastError();
      return fail(MC_EHIP, "mc_measure_noalloc: hipGetDeviceProperties failed");
    }
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
      return fail(MC_ENODEV, std::string("mc_measure_noalloc: current device is ") + prop.gcnArchName + ", not gfx950");
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      return fail(MC_EHIP, std::string("mc_measure_noalloc: hipGetSymbolAddress(scratch) failed: ") +
                               hipGetErrorString(e));
    }
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && lds > 0)
      s.lds_limit = (size_t)lds;
    (void)hipGetLastError();
    s.cu_count = prop.multiProcessorCount;
    (void)mc_slab_prepare(s.lds_limit);
    s.ptr = static_cast<double *>(p);
    s.ready = true;
  }
  *out = s;
  return MC_OK;
}

__global__ void k_noalloc_state(MCDevState *st, float T, float elecE) {
  st->T = T;
  st->elecE = elecE;
  st->ang_var = 0.0f;
  st->sweep = 0u;
  st->acc_count = 0u;
  st->blocks_done = 0u;
  st->last_acc = 0u;
  st->meas_done = 0u;
}

}  // namespace

extern "C" int mc_measure_noalloc(const mc_params *p, const float *nx, const float *ny, const float *nz, double *out,
                                  void *stream) {
  if (!p || !nx || !ny || !nz || !out) return fail(MC_EINVAL, "mc_measure_noalloc: null argument");
  if (p->nx < 1 || p->ny < 1 || p->nz < 1) return fail(MC_EINVAL, "mc_measure_noalloc: lattice dimensions must be >= 1");
  const long long ntl = (long long)p->nx * p->ny * p->nz;
  if (ntl >= (1LL << 31)) return fail(MC_EINVAL, "mc_measure_noalloc: lattice of 2^31 sites or more is not supported");
  if (p->potential < MC_LL || p->potential > MC_PEAR) return fail(MC_EINVAL, "mc_measure_noalloc: unknown potential");
  if (p->nx % 4 != 0 || p->bound[0] != MC_PERIODIC || p->bound[1] != MC_PERIODIC || p->ny % 2 != 0)
    return fail(MC_EINVAL, "mc_measure_noalloc: needs Nx a multiple of 4, periodic x and y, even Ny");
  if (p->bound[2] != MC_PERIODIC && p->bound[2] != MC_FREE)
    return fail(MC_EINVAL, "mc_measure_noalloc: boundary must be MC_FREE or MC_PERIODIC");
  if (p->bound[2] == MC_PERIODIC && (p->nz % 2 != 0 || p->n_surf != 0))
    return fail(MC_EINVAL, "mc_measure_noalloc: a periodic z needs even Nz and no surfaces");
  if (p->bound[2] == MC_FREE && (p->nz < 2 || (p->n_surf != 0 && p->n_surf < 2) || p->n_surf > MC_MAX_SURF))
    return fail(MC_EINVAL, "mc_measure_noalloc: a free z takes no surfaces or the slab faces (surfaces 0 and 1)");
  for (int s = 0; s < p->n_surf; ++s) {
    if (p->surf_kind[s] < MC_HOMEOTROPIC || p->surf_kind[s] > MC_STRONG_GHRL)
      return fail(MC_EINVAL, "mc_measure_noalloc: unknown anchoring kind");
    const int k = p->surf_kind[s];
    const float *d = p->surf_dir[s];
    const float d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    if ((k == MC_RP || k == MC_STRONG || k == MC_RP_GHRL || k == MC_STRONG_GHRL) && !(fabsf(d2 - 1.0f) <= 1e-3f))
      return fail(MC_EINVAL, "mc_measure_noalloc: surf_dir must be a unit vector for the rp and strong kinds");
  }

  DeviceSlot slot;
  int rc = device_slot(&slot);
  if (rc) return rc;

  MCLaunch L{};
  mc_to_dev_params(p, &L.P);
  L.P.nt = (int)ntl;
  L.P.valid = (int)ntl;
  L.P.zfree = p->bound[2] == MC_FREE ? 1 : 0;
  L.P.zsurf = (p->bound[2] == MC_FREE && p->n_surf >= 2) ? 1 : 0;
  L.stream = static_cast<hipStream_t>(stream);
  L.nx = const_cast<float *>(nx);
  L.ny = const_cast<float *>(ny);
  L.nz = const_cast<float *>(nz);
  L.st = reinterpret_cast<MCDevState *>(slot.ptr);
  double *partials = slot.ptr + kStateDoubles;

  // The partial-sum grid mc_measure uses for this lattice (one row per
  // workgroup), which must fit the scratch array.
  long long grid = (ntl + 255) / 256;
  if (grid > 8LL * slot.cu_count) grid = 8LL * slot.cu_count;
  if (grid < 1) grid = 1;
  if (grid > kMaxParts)
    return fail(MC_EINVAL, "mc_measure_noalloc: the lattice needs " + std::to_string(grid) +
                               " rows of partial sums, above the scratch maximum of " + std::to_string(kMaxParts));
  MCSlabShape shape{0, 0, 0};
  const bool slab = mc_measure_slab_shape(L.P, slot.lds_limit, slot.cu_count, (int)grid, &shape);

  k_noalloc_state<<<1, 1, 0, L.stream>>>(L.st, p->T, p->elecE);
  if (slab)
    mc_launch_measure_slab(L, shape, partials, out, nullptr);
  else
    mc_launch_measure_x4(L, (int)grid, grid % 8 == 0 ? 1 : 0, partials, out, nullptr);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(MC_EHIP, std::string("mc_measure_noalloc: launch failed: ") + hipGetErrorString(e));
  return MC_OK;
}
