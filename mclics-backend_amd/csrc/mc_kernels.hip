// Sweep, measurement and transfer kernels of the checkerboard Metropolis
// engine (gfx950, wave64).
//
// Lattice layout in HBM: three FP32 planes nx[Nt], ny[Nt], nz[Nt] indexed
// s = i + Nx*(j + Ny*k) (the reference's site order, AoS split into SoA so a
// wave's loads of one component are contiguous).  Non-bulk lattices add a
// uint8 point-type plane and SoA surface normals; bulk lattices read neither.
//
// One pass updates one non-interacting site set: a 2x2x2 sublattice tick
// (reference order, src/evolve.cpp:90-113 / src/evolve.cu:123-137) or one
// checkerboard colour.  Passes are grid-stride loops over a few thousand
// workgroups so the per-workgroup acceptance atomics stay cheap; the last
// workgroup of the final pass of a sweep adapts ang_var on the device
// (src/evolve.cu:378-384) so sweeps never wait for the host.
#include "mc_kernels.h"

namespace {

constexpr int kBlock = MCD_BLOCK;

// Neighbour indices in the reference order +x,-x,+y,-y,+z,-z and whether each
// bond exists: periodic axes wrap, free axes drop out-of-range neighbours
// (src/potential.cpp:44-54), inactive neighbours drop out (pt==0).
template <bool GEN>
__device__ __forceinline__ void mcd_neighbours(const MCDevParams &P, const uint8_t *pt, int i, int j, int k,
                                               int nb[6], bool on[6]) {
  const int nx = P.nx, ny = P.ny, nz = P.nz;
  const int sxy = nx * ny;
  const int base = i + nx * (j + ny * k);
  int ip = i + 1, im = i - 1, jp = j + 1, jm = j - 1, kp = k + 1, km = k - 1;
  on[0] = on[1] = on[2] = on[3] = on[4] = on[5] = true;
  if (ip == nx) { if (P.per[0]) ip = 0; else on[0] = false; }
  if (im < 0) { if (P.per[0]) im = nx - 1; else on[1] = false; }
  if (jp == ny) { if (P.per[1]) jp = 0; else on[2] = false; }
  if (jm < 0) { if (P.per[1]) jm = ny - 1; else on[3] = false; }
  if (kp == nz) { if (P.per[2]) kp = 0; else on[4] = false; }
  if (km < 0) { if (P.per[2]) km = nz - 1; else on[5] = false; }
  nb[0] = base + (ip - i);
  nb[1] = base + (im - i);
  nb[2] = base + nx * (jp - j);
  nb[3] = base + nx * (jm - j);
  nb[4] = base + sxy * (kp - k);
  nb[5] = base + sxy * (km - k);
  if (GEN) {
#pragma unroll
    for (int b = 0; b < 6; ++b)
      if (on[b]) on[b] = pt[nb[b]] != 0;
  }
}

// Old (and optionally new) site energy of site s, the reference's
// latice_Potential (src/geometry_bulk.cpp:63-72, src/geometry_slab.cpp:69-83).
template <int POT, bool GEN, bool TWO>
__device__ __forceinline__ void mcd_site_energy(const MCDevParams &P, double cA, const float *__restrict__ nx,
                                                const float *__restrict__ ny, const float *__restrict__ nz,
                                                const uint8_t *__restrict__ pt, const float *__restrict__ sx,
                                                const float *__restrict__ sy, const float *__restrict__ sz, float elecE,
                                                int i, int j, int k, int s, float o0, float o1, float o2, float q0,
                                                float q1, float q2, float &Eo, float &En) {
  int nb[6];
  bool on[6];
  mcd_neighbours<GEN>(P, pt, i, j, k, nb, on);
  double so = 0.0, sn = 0.0;
#pragma unroll
  for (int b = 0; b < 6; ++b) {
    if (!on[b]) continue;
    const float m0 = nx[nb[b]], m1 = ny[nb[b]], m2 = nz[nb[b]];
    so += (double)mcd_bond<POT>(P, cA, b, o0, o1, o2, m0, m1, m2);
    if (TWO) sn += (double)mcd_bond<POT>(P, cA, b, q0, q1, q2, m0, m1, m2);
  }
  Eo = (float)so;
  En = (float)sn;
  if (GEN) {
    const int p = pt[s];
    if (p > 1 && P.n_surf > 0) {
      const float s0 = sx[s], s1 = sy[s], s2 = sz[s];
      Eo += mcd_surface(P, p - 2, o0, o1, o2, s0, s1, s2);
      if (TWO) En += mcd_surface(P, p - 2, q0, q1, q2, s0, s1, s2);
    }
  }
  if (P.elecA != 0.0f) {
    Eo += mcd_field(P, elecE, o0, o1, o2);
    if (TWO) En += mcd_field(P, elecE, q0, q1, q2);
  }
}

// Production pass: counter-based proposal, fused old/new energy, accept.
template <int POT, bool GEN, bool RECORD>
__global__ __launch_bounds__(kBlock) void k_sweep_pass(MCDevParams P, float *__restrict__ nx, float *__restrict__ ny,
                                                       float *__restrict__ nz, const uint8_t *__restrict__ pt,
                                                       const float *__restrict__ sx, const float *__restrict__ sy,
                                                       const float *__restrict__ sz, MCDevState *st, int set,
                                                       int total, int finalize, MCRecord rec) {
  const float T = st->T, elecE = st->elecE, av = st->ang_var;
  const uint32_t sweep = st->sweep;
  const double cA = (double)(-P.A) * 1.5;
  uint32_t acc = 0;
  for (int t = blockIdx.x * kBlock + threadIdx.x; t < total; t += gridDim.x * kBlock) {
    int i, j, k;
    if (!mcd_set_site(P, set, t, i, j, k)) continue;
    const int s = i + P.nx * (j + P.ny * k);
    if (GEN && pt[s] == 0) continue;
    const MCU4 r = mcd_philox(MCU4{(uint32_t)s, sweep, P.stream, 0x4D43u}, P.seed_lo, P.seed_hi);
    const float o0 = nx[s], o1 = ny[s], o2 = nz[s];
    // Proposal of src/evolve.cu:143-163: rotate by pi*va about x, y or z
    // (thresholds 0.333/0.666 on a 24-bit uniform, applied on its integer
    // mantissa) and renormalise.  v_sin/v_cos take revolutions, so
    // sin(pi*va) = sin(2*pi*(va/2)) needs no range reduction.
    const float u0 = mcd_u01(r.x);
    const uint32_t m1 = r.y >> 8;
    const float va = (2.0f * u0 - 1.0f) * av;
    const float hrev = 0.5f * va;
    const float sn = __builtin_amdgcn_sinf(hrev), cs = __builtin_amdgcn_cosf(hrev);
    float q0, q1, q2;
    if (m1 < 5586813u) {
      q0 = o0;
      q1 = o1 * cs + o2 * sn;
      q2 = o2 * cs - o1 * sn;
    } else if (m1 < 11173626u) {
      q0 = o0 * cs - o2 * sn;
      q1 = o1;
      q2 = o2 * cs + o0 * sn;
    } else {
      q0 = o0 * cs + o1 * sn;
      q1 = o1 * cs - o0 * sn;
      q2 = o2;
    }
    const float inv = __builtin_amdgcn_rsqf(q0 * q0 + q1 * q1 + q2 * q2);
    q0 *= inv;
    q1 *= inv;
    q2 *= inv;
    const float u = mcd_u01(r.z);
    float Eo, En;
    mcd_site_energy<POT, GEN, true>(P, cA, nx, ny, nz, pt, sx, sy, sz, elecE, i, j, k, s, o0, o1, o2, q0, q1, q2,
                                    Eo, En);
    const bool a = mcd_accept(En, Eo, T, u);
    if (a) {
      nx[s] = q0;
      ny[s] = q1;
      nz[s] = q2;
    }
    acc += a ? 1u : 0u;
    if (RECORD) {
      rec.u0[s] = u0;
      rec.u1[s] = mcd_u01(r.y);
      rec.u2[s] = u;
      rec.px[s] = q0;
      rec.py[s] = q1;
      rec.pz[s] = q2;
      rec.acc[s] = a ? 1 : 0;
      rec.eo[s] = Eo;
      rec.en[s] = En;
    }
  }
  const uint32_t bsum = mcd_block_sum(acc);
  mcd_finish_sweep(P, st, bsum, finalize != 0);
}

// Parity pass with injected proposals (AoS) and uniforms.
template <int POT, bool GEN>
__global__ __launch_bounds__(kBlock) void k_pass_inject(MCDevParams P, float *__restrict__ nx, float *__restrict__ ny,
                                                        float *__restrict__ nz, const uint8_t *__restrict__ pt,
                                                        const float *__restrict__ sx, const float *__restrict__ sy,
                                                        const float *__restrict__ sz, const MCDevState *st, int set,
                                                        int total, const float *__restrict__ prop,
                                                        const float *__restrict__ uin, uint8_t *__restrict__ acc,
                                                        float *__restrict__ eo, float *__restrict__ en) {
  const float T = st->T, elecE = st->elecE;
  const double cA = (double)(-P.A) * 1.5;
  for (int t = blockIdx.x * kBlock + threadIdx.x; t < total; t += gridDim.x * kBlock) {
    int i, j, k;
    if (!mcd_set_site(P, set, t, i, j, k)) continue;
    const int s = i + P.nx * (j + P.ny * k);
    if (GEN && pt[s] == 0) continue;
    const float o0 = nx[s], o1 = ny[s], o2 = nz[s];
    const float q0 = prop[3 * (size_t)s], q1 = prop[3 * (size_t)s + 1], q2 = prop[3 * (size_t)s + 2];
    float Eo, En;
    mcd_site_energy<POT, GEN, true>(P, cA, nx, ny, nz, pt, sx, sy, sz, elecE, i, j, k, s, o0, o1, o2, q0, q1, q2,
                                    Eo, En);
    const bool a = mcd_accept(En, Eo, T, uin[s]);
    if (a) {
      nx[s] = q0;
      ny[s] = q1;
      nz[s] = q2;
    }
    acc[s] = a ? 1 : 0;
    eo[s] = Eo;
    en[s] = En;
  }
}

// Site energies of every site (0 where pt==0).
template <int POT, bool GEN>
__global__ __launch_bounds__(kBlock) void k_site_energy(MCDevParams P, const float *__restrict__ nx,
                                                        const float *__restrict__ ny, const float *__restrict__ nz,
                                                        const uint8_t *__restrict__ pt, const float *__restrict__ sx,
                                                        const float *__restrict__ sy, const float *__restrict__ sz,
                                                        const MCDevState *st, float *__restrict__ out) {
  const float elecE = st->elecE;
  const double cA = (double)(-P.A) * 1.5;
  for (int s = blockIdx.x * kBlock + threadIdx.x; s < P.nt; s += gridDim.x * kBlock) {
    if (GEN && pt[s] == 0) {
      out[s] = 0.0f;
      continue;
    }
    const int i = s % P.nx, j = (s / P.nx) % P.ny, k = s / (P.nx * P.ny);
    float Eo, En;
    mcd_site_energy<POT, GEN, false>(P, cA, nx, ny, nz, pt, sx, sy, sz, elecE, i, j, k, s, nx[s], ny[s], nz[s], 0.f,
                                     0.f, 0.f, Eo, En);
    out[s] = Eo;
  }
}

// FP64 partial sums per workgroup: {E, Qxx, Qxy, Qxz, Qyy, Qyz, Qzz, n}
// with Q accumulated as sum(3 n_a n_b - delta_ab) (src/parameter_order.cpp:30-60).
template <int POT, bool GEN>
__global__ __launch_bounds__(kBlock) void k_measure(MCDevParams P, const float *__restrict__ nx,
                                                    const float *__restrict__ ny, const float *__restrict__ nz,
                                                    const uint8_t *__restrict__ pt, const float *__restrict__ sx,
                                                    const float *__restrict__ sy, const float *__restrict__ sz,
                                                    const MCDevState *st, double *__restrict__ partials) {
  const float elecE = st->elecE;
  const double cA = (double)(-P.A) * 1.5;
  double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int s = blockIdx.x * kBlock + threadIdx.x; s < P.nt; s += gridDim.x * kBlock) {
    if (GEN && pt[s] == 0) continue;
    const int i = s % P.nx, j = (s / P.nx) % P.ny, k = s / (P.nx * P.ny);
    const float o0 = nx[s], o1 = ny[s], o2 = nz[s];
    float Eo, En;
    mcd_site_energy<POT, GEN, false>(P, cA, nx, ny, nz, pt, sx, sy, sz, elecE, i, j, k, s, o0, o1, o2, 0.f, 0.f,
                                     0.f, Eo, En);
    const double x = o0, y = o1, z = o2;
    a[0] += (double)Eo;
    a[1] += 3.0 * x * x - 1.0;
    a[2] += 3.0 * x * y;
    a[3] += 3.0 * x * z;
    a[4] += 3.0 * y * y - 1.0;
    a[5] += 3.0 * y * z;
    a[6] += 3.0 * z * z - 1.0;
    a[7] += 1.0;
  }
  __shared__ double red[8][kBlock / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    double v = a[c];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if (lane == 0) red[c][wid] = v;
  }
  __syncthreads();
  if (threadIdx.x < 8) {
    double v = 0;
    for (int w = 0; w < kBlock / 64; ++w) v += red[threadIdx.x][w];
    partials[(size_t)blockIdx.x * 8 + threadIdx.x] = v;
  }
}


// Local order parameter of every site, src/parameter_order.cpp:62-113: the
// Q tensor of the site and its active von Neumann neighbours (wrapped on a
// periodic axis, dropped beyond a free face), in FP64, normalised by
// 2 * points, and S = its largest eigenvalue; NaN where pt == 0.  Neighbour
// order as the reference's di/dj/dk loops.
__global__ __launch_bounds__(kBlock) void k_local_S(MCDevParams P, const float *__restrict__ nx,
                                                    const float *__restrict__ ny, const float *__restrict__ nz,
                                                    const uint8_t *__restrict__ pt, float *__restrict__ out) {
  const int off[7][3] = {{-1, 0, 0}, {0, -1, 0}, {0, 0, -1}, {0, 0, 0}, {0, 0, 1}, {0, 1, 0}, {1, 0, 0}};
  const int dims[3] = {P.nx, P.ny, P.nz};
  for (int s = blockIdx.x * kBlock + threadIdx.x; s < P.nt; s += gridDim.x * kBlock) {
    if (pt && pt[s] == 0) {
      out[s] = __int_as_float(0x7fc00000);  // NaN
      continue;
    }
    const int c0[3] = {s % P.nx, (s / P.nx) % P.ny, s / (P.nx * P.ny)};
    double q[6] = {0, 0, 0, 0, 0, 0};
    int points = 0;
#pragma unroll
    for (int o = 0; o < 7; ++o) {
      int c[3];
      bool ok = true;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        c[a] = c0[a] + off[o][a];
        if (c[a] < 0 || c[a] >= dims[a]) {
          if (P.per[a]) c[a] = c[a] < 0 ? c[a] + dims[a] : c[a] - dims[a];
          else ok = false;
        }
      }
      if (!ok) continue;
      const int t = c[0] + P.nx * (c[1] + P.ny * c[2]);
      if (pt && pt[t] == 0) continue;
      const double x = nx[t], y = ny[t], z = nz[t];
      q[0] += 3.0 * (x * x);
      q[1] += 3.0 * (x * y);
      q[2] += 3.0 * (x * z);
      q[3] += 3.0 * (y * y);
      q[4] += 3.0 * (y * z);
      q[5] += 3.0 * (z * z);
      q[0] -= 1.0;
      q[3] -= 1.0;
      q[5] -= 1.0;
      ++points;
    }
    const double norm = 2.0 * points;
#pragma unroll
    for (int c = 0; c < 6; ++c) q[c] /= norm;
    out[s] = (float)mcd_sym3_max(q);
  }
}

// Second stage, one workgroup, fixed summation order (deterministic):
// out = {E_mean, S, Qxx, Qxy, Qxz, Qyy, Qyz, Qzz, n_active}.
__global__ __launch_bounds__(kBlock) void k_measure_final(const double *__restrict__ partials, int nparts,
                                                          double *__restrict__ out, double *__restrict__ series) {
  __shared__ double red[8][kBlock];
  double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int p = threadIdx.x; p < nparts; p += kBlock)
#pragma unroll
    for (int c = 0; c < 8; ++c) a[c] += partials[(size_t)p * 8 + c];
#pragma unroll
  for (int c = 0; c < 8; ++c) red[c][threadIdx.x] = a[c];
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w)
#pragma unroll
      for (int c = 0; c < 8; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double n = red[7][0];
    double q[6];
    for (int c = 0; c < 6; ++c) q[c] = n > 0 ? red[c + 1][0] / (2.0 * n) : 0.0;
    out[0] = n > 0 ? red[0][0] / n : 0.0;
    out[1] = mcd_sym3_max(q);
    for (int c = 0; c < 6; ++c) out[2 + c] = q[c];
    out[8] = n;
    if (series) {
      series[0] = out[0];
      series[1] = out[1];
    }
  }
}

__global__ void k_aos_to_soa(const float *__restrict__ aos, float *__restrict__ x, float *__restrict__ y,
                             float *__restrict__ z, int n) {
  for (int s = blockIdx.x * kBlock + threadIdx.x; s < n; s += gridDim.x * kBlock) {
    x[s] = aos[3 * (size_t)s];
    y[s] = aos[3 * (size_t)s + 1];
    z[s] = aos[3 * (size_t)s + 2];
  }
}

__global__ void k_soa_to_aos(const float *__restrict__ x, const float *__restrict__ y, const float *__restrict__ z,
                             float *__restrict__ aos, int n) {
  for (int s = blockIdx.x * kBlock + threadIdx.x; s < n; s += gridDim.x * kBlock) {
    aos[3 * (size_t)s] = x[s];
    aos[3 * (size_t)s + 1] = y[s];
    aos[3 * (size_t)s + 2] = z[s];
  }
}

// One workgroup per replica sums its tiles in tile order (deterministic):
// out[r*8..] = {E_mean, S, q6}; series[r*2..] = {E, S} when given.
__global__ __launch_bounds__(kBlock) void k_batch_measure_final(const int *__restrict__ rep_tile0,
                                                                const int *__restrict__ rep_ntiles,
                                                                const double *__restrict__ partials,
                                                                double *__restrict__ out, double *__restrict__ series) {
  __shared__ double red[8][kBlock];
  const int r = blockIdx.x;
  double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int p = threadIdx.x; p < rep_ntiles[r]; p += kBlock)
#pragma unroll
    for (int c = 0; c < 8; ++c) a[c] += partials[(size_t)(rep_tile0[r] + p) * 8 + c];
#pragma unroll
  for (int c = 0; c < 8; ++c) red[c][threadIdx.x] = a[c];
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w)
#pragma unroll
      for (int c = 0; c < 8; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double n = red[7][0];
    double q[6];
    for (int c = 0; c < 6; ++c) q[c] = red[c + 1][0] / (2.0 * n);
    double *o = out + (size_t)r * 8;
    o[0] = red[0][0] / n;
    o[1] = mcd_sym3_max(q);
    for (int c = 0; c < 6; ++c) o[2 + c] = q[c];
    if (series) {
      series[2 * (size_t)r] = o[0];
      series[2 * (size_t)r + 1] = o[1];
    }
  }
}

template <int POT, bool GEN>
void launch_sweep_pass_t(const MCLaunch &L, int set, int total, int grid, int finalize, const MCRecord *rec) {
  if (rec)
    k_sweep_pass<POT, GEN, true><<<grid, kBlock, 0, L.stream>>>(L.P, L.nx, L.ny, L.nz, L.pt, L.sx, L.sy, L.sz, L.st,
                                                                set, total, finalize, *rec);
  else
    k_sweep_pass<POT, GEN, false><<<grid, kBlock, 0, L.stream>>>(L.P, L.nx, L.ny, L.nz, L.pt, L.sx, L.sy, L.sz,
                                                                 L.st, set, total, finalize, MCRecord{});
}

}  // namespace

// Number of slots of a site set (threads the pass maps).
int mc_set_slots(const MCDevParams &P, int set) { return mcd_set_slots(P, set); }

void mc_launch_sweep_pass(const MCLaunch &L, int set, int grid, int finalize, const MCRecord *rec) {
  const int total = mc_set_slots(L.P, set);
  const bool gen = L.pt != nullptr;
  switch (L.P.potential) {
    case MCD_GHRL:
      gen ? launch_sweep_pass_t<MCD_GHRL, true>(L, set, total, grid, finalize, rec)
          : launch_sweep_pass_t<MCD_GHRL, false>(L, set, total, grid, finalize, rec);
      break;
    case MCD_PEAR:
      gen ? launch_sweep_pass_t<MCD_PEAR, true>(L, set, total, grid, finalize, rec)
          : launch_sweep_pass_t<MCD_PEAR, false>(L, set, total, grid, finalize, rec);
      break;
    default:
      gen ? launch_sweep_pass_t<MCD_LL, true>(L, set, total, grid, finalize, rec)
          : launch_sweep_pass_t<MCD_LL, false>(L, set, total, grid, finalize, rec);
      break;
  }
}

#define MC_DISPATCH(KERNEL, GRID, ...)                                                        \
  do {                                                                                       \
    const bool gen_ = L.pt != nullptr;                                                       \
    switch (L.P.potential) {                                                                 \
      case MCD_GHRL:                                                                         \
        if (gen_) KERNEL<MCD_GHRL, true><<<GRID, kBlock, 0, L.stream>>>(__VA_ARGS__);        \
        else KERNEL<MCD_GHRL, false><<<GRID, kBlock, 0, L.stream>>>(__VA_ARGS__);            \
        break;                                                                               \
      case MCD_PEAR:                                                                         \
        if (gen_) KERNEL<MCD_PEAR, true><<<GRID, kBlock, 0, L.stream>>>(__VA_ARGS__);        \
        else KERNEL<MCD_PEAR, false><<<GRID, kBlock, 0, L.stream>>>(__VA_ARGS__);            \
        break;                                                                               \
      default:                                                                               \
        if (gen_) KERNEL<MCD_LL, true><<<GRID, kBlock, 0, L.stream>>>(__VA_ARGS__);          \
        else KERNEL<MCD_LL, false><<<GRID, kBlock, 0, L.stream>>>(__VA_ARGS__);              \
        break;                                                                               \
    }                                                                                        \
  } while (0)

void mc_launch_pass_inject(const MCLaunch &L, int set, int grid, const float *prop, const float *u, uint8_t *acc,
                           float *eo, float *en) {
  const int total = mc_set_slots(L.P, set);
  MC_DISPATCH(k_pass_inject, grid, L.P, L.nx, L.ny, L.nz, L.pt, L.sx, L.sy, L.sz, L.st, set, total, prop, u, acc, eo,
              en);
}

void mc_launch_site_energy(const MCLaunch &L, int grid, float *out) {
  MC_DISPATCH(k_site_energy, grid, L.P, L.nx, L.ny, L.nz, L.pt, L.sx, L.sy, L.sz, L.st, out);
}

void mc_launch_measure_final(hipStream_t stream, const double *partials, int nparts, double *out, double *series) {
  k_measure_final<<<1, kBlock, 0, stream>>>(partials, nparts, out, series);
}

void mc_launch_measure(const MCLaunch &L, int grid, double *partials, double *out, double *series) {
  MC_DISPATCH(k_measure, grid, L.P, L.nx, L.ny, L.nz, L.pt, L.sx, L.sy, L.sz, L.st, partials);
  k_measure_final<<<1, kBlock, 0, L.stream>>>(partials, grid, out, series);
}

void mc_launch_local_S(const MCLaunch &L, int grid, float *out) {
  k_local_S<<<grid, kBlock, 0, L.stream>>>(L.P, L.nx, L.ny, L.nz, L.pt, out);
}

void mc_launch_aos_to_soa(hipStream_t s, const float *aos, float *x, float *y, float *z, int n, int grid) {
  k_aos_to_soa<<<grid, kBlock, 0, s>>>(aos, x, y, z, n);
}

void mc_launch_soa_to_aos(hipStream_t s, const float *x, const float *y, const float *z, float *aos, int n,
                          int grid) {
  k_soa_to_aos<<<grid, kBlock, 0, s>>>(x, y, z, aos, n);
}

void mc_launch_batch_measure(const MCBatchLaunch &B, int grid, double *tile_partials, double *out, double *series) {
  mc_launch_batch_measure_tiles(B, grid, tile_partials);
  mc_launch_batch_measure_final(B, B.rep_tile0, B.rep_ntiles, tile_partials, out, series);
}

void mc_launch_batch_measure_final(const MCBatchLaunch &B, const int *first, const int *count, const double *partials,
                                   double *out, double *series) {
  k_batch_measure_final<<<B.nrep, kBlock, 0, B.stream>>>(first, count, partials, out, series);
}
