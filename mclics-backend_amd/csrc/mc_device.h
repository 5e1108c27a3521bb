// Device-side physics of the checkerboard Metropolis sweep (gfx950).
//
// Every energy expression reproduces the reference CPU path operation by
// operation: FP32 dot products evaluated ((a0*b0 + a1*b1) + a2*b2) with no
// FMA contraction (this translation unit is compiled with -ffp-contract=off),
// the double literals of the LL/GHRL/pear bonds kept as FP64 arithmetic, and
// the six-bond sum accumulated in FP64 and rounded once to FP32
// (src/geometry.cpp:19-72).  CDNA4 runs FP64 FMA/mul at the same issue rate as
// unpacked FP32, so keeping the reference's mixed precision costs conversion
// instructions only.
#pragma once
#include <stddef.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mc_expf.h"

#define MCD_MAX_SURF 8
#define MCD_BLOCK 256  // threads per workgroup of every sweep kernel

enum { MCD_LL = 0, MCD_GHRL = 1, MCD_PEAR = 2 };
enum {
  MCD_HOMEOTROPIC = 0,
  MCD_RP = 1,
  MCD_FG = 2,
  MCD_STRONG = 3,
  MCD_HOMEOTROPIC_GHRL = 4,
  MCD_RP_GHRL = 5,
  MCD_FG_GHRL = 6,
  MCD_STRONG_GHRL = 7
};

// Immutable per-handle parameters, passed by value as a kernel argument.
// Exact unsigned division by a divisor fixed for a launch, with one
// multiply-high (Granlund & Montgomery, PLDI 1994, round-up method): for
// 32-bit n, q = (t + ((n - t) >> 1)) >> (l - 1) with t = mulhi(m, n),
// l = ceil(log2 d), m = floor(2^32 (2^l - d) / d) + 1; d = 1 gives q = n.
struct MCDiv {
  uint32_t m, sh, d;
};

__host__ __device__ inline MCDiv mcd_div_make(uint32_t d) {
  MCDiv r{0u, 0u, d};
  if (d <= 1u) return r;
  uint32_t l = 0;
  while ((1ull << l) < d) ++l;
  r.m = (uint32_t)(((1ull << 32) * ((1ull << l) - d)) / d + 1ull);
  r.sh = l - 1;
  return r;
}

__device__ __forceinline__ uint32_t mcd_div(uint32_t n, const MCDiv &D) {
  if (D.d <= 1u) return n;
  const uint32_t t = __umulhi(D.m, n);
  return (t + ((n - t) >> 1)) >> D.sh;
}

struct MCDevParams {
  int nx, ny, nz;
  int per[3];  // 1 periodic, 0 free
  int potential;
  float A, B1, C;
  float rho, lambda, mu, nu, sigma;
  float eX, eY, eZ, elecA;
  int n_surf;
  int surf_kind[MCD_MAX_SURF];
  float surf_W[MCD_MAX_SURF];
  float surf_dir[MCD_MAX_SURF][3];
  uint32_t seed_lo, seed_hi, stream;
  int nt;
  int valid;
  // dividers of the site-set maps: (nx+1)/2, ny, Ni = (nx+1)/2, Nj = (ny+1)/2, Ni*Nj
  MCDiv div_hx, div_ny, div_nj, div_ninj;
  // MCD_NSUB sweep sub-counters, MCD_SUB_STRIDE words apart (mcd_finish_sweep)
  unsigned long long *ctr;
  // Slab geometry on the fast kernels (checked against pt and ns at create):
  // zfree: z is free (x, y periodic, every site active); zsurf: the planes
  // k = 0 and k = Nz-1 carry surfaces 0 and 1 with normals (0, 0, -1) and
  // (0, 0, +1), the layout of src/geometry_slab.cpp:47-67.
  int zfree, zsurf;
};

// Sub-counters of the sweep bookkeeping: same-address device atomics
// serialise (about 11.5 ns each on MI355X, so 2048 workgroups counting into
// one word cost ~24 us per pass), while atomics on different 128-byte lines
// proceed in parallel.
#define MCD_NSUB 64
#define MCD_SUB_STRIDE 16

// Mutable sweep state in device memory, so captured graphs replay with the
// current temperature, field and proposal width.
struct MCDevState {
  float T;
  float elecE;
  float ang_var;
  uint32_t sweep;        // index of the next sweep (RNG counter word 1)
  uint32_t acc_count;    // accepted moves of the sweep in progress (8-byte aligned: low word of
  uint32_t blocks_done;  //   one 64-bit counter with blocks_done, last-block detection of the final pass)
  uint32_t last_acc;     // accepted moves of the last completed sweep
  uint32_t meas_done;    // spare (zero)
};
static_assert(offsetof(MCDevState, acc_count) % 8 == 0 && offsetof(MCDevState, blocks_done) == offsetof(MCDevState, acc_count) + 4,
              "acc_count/blocks_done form one 64-bit counter");

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11): two 32x32->64 products per round,
// Weyl key increments between rounds.  (uint64_t)a*b lowers to
// v_mad_u64_u32, one full-rate instruction for both product halves, and each
// three-way XOR is one CDNA4 v_bitop3_b32 (truth table 0x96 = a ^ b ^ c),
// which the compiler does not form by itself from two XORs.
struct MCU4 {
  uint32_t x, y, z, w;
};

__device__ __forceinline__ uint32_t mcd_xor3(uint32_t a, uint32_t b, uint32_t c) {
  return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

__device__ __forceinline__ MCU4 mcd_philox(MCU4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
    MCU4 n;
    n.x = mcd_xor3((uint32_t)(p1 >> 32), c.y, k0);
    n.y = (uint32_t)p1;
    n.z = mcd_xor3((uint32_t)(p0 >> 32), c.w, k1);
    n.w = (uint32_t)p0;
    c = n;
  }
  return c;
}

// 24-bit uniform on [0,1), exact in FP32.
__device__ __forceinline__ float mcd_u01(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }

// ---------------------------------------------------------------------------
// Site sets of a pass.

// Maps a linear slot of a site set to (i, j, k); returns false for slots
// outside the lattice.
__device__ __forceinline__ bool mcd_set_site(const MCDevParams &P, int set, int t, int &i, int &j, int &k) {
  if (set >= 8) {
    const int c = set - 8;
    const int row = (int)mcd_div((uint32_t)t, P.div_hx);
    const int h = t - row * (int)P.div_hx.d;
    k = (int)mcd_div((uint32_t)row, P.div_ny);
    j = row - k * P.ny;
    i = 2 * h + ((j + k + c) & 1);
    return i < P.nx && k < P.nz;
  }
  const int di = set & 1, dj = (set >> 1) & 1, dk = set >> 2;
  const int r1 = (int)mcd_div((uint32_t)t, P.div_hx);  // t / Ni (Ni = hx)
  const int kk = (int)mcd_div((uint32_t)t, P.div_ninj);
  i = di + 2 * (t - r1 * (int)P.div_hx.d);
  j = dj + 2 * (r1 - (int)mcd_div((uint32_t)r1, P.div_nj) * (int)P.div_nj.d);
  k = dk + 2 * kk;
  return i < P.nx && j < P.ny && k < P.nz;
}

// Number of slots of a site set (threads a pass maps).
__device__ __host__ __forceinline__ int mcd_set_slots(const MCDevParams &P, int set) {
  if (set >= 8) return ((P.nx + 1) / 2) * P.ny * P.nz;
  return ((P.nx + 1) / 2) * ((P.ny + 1) / 2) * ((P.nz + 1) / 2);
}

// ---------------------------------------------------------------------------
// Energy terms.

__device__ __forceinline__ float mcd_dot(float a0, float a1, float a2, float b0, float b1, float b2) {
  float t = a0 * b0;
  t = t + a1 * b1;
  t = t + a2 * b2;
  return t;
}

// src/potential.cpp:17-20: -A*1.5*nij*nij evaluated in double.
__device__ __forceinline__ float mcd_bond_ll(double cA, float nij) {
  double v = cA * (double)nij;
  v = v * (double)nij;
  return (float)v;
}

// src/potential.cpp:22-38 (CPU form: E1 built in double, em*(...-1.0/9) in
// double, the sum in double).
__device__ __forceinline__ float mcd_bond_ghrl(const MCDevParams &P, float n0, float n1, float n2, float m0,
                                               float m1, float m2, float r0, float r1, float r2) {
  float ai = mcd_dot(n0, n1, n2, r0, r1, r2);
  float aj = mcd_dot(m0, m1, m2, r0, r1, r2);
  float nij = mcd_dot(n0, n1, n2, m0, m1, m2);
  float pij = 1.5f * nij * nij - 0.5f;
  float c0 = n2 * m1 - n1 * m2;
  float c1 = n0 * m2 - n2 * m0;
  float c2 = n1 * m0 - n0 * m1;
  float cross = c0 * r0;
  cross = cross + c1 * r1;
  cross = cross + c2 * r2;
  double e1d = 1.5 * (double)ai * (double)ai;
  e1d = e1d + 1.5 * (double)aj * (double)aj;
  e1d = e1d - 1.0;
  float E1 = (float)e1d;
  float t1 = E1 * (P.rho * pij + P.lambda);
  float aan = ai * aj * nij;
  double t2 = (double)P.mu * ((double)aan - 1.0 / 9);
  double s = (double)t1 + t2;
  s = s + (double)(P.nu * pij);
  float sg = nij > 0 ? 1.0f : -1.0f;
  s = s + (double)(P.sigma * sg * cross);
  return (float)s;
}

// src/potential.cpp:10-15.
__device__ __forceinline__ float mcd_bond_pear(const MCDevParams &P, float n0, float n1, float n2, float m0,
                                               float m1, float m2, float r0, float r1, float r2) {
  float nij = mcd_dot(n0, n1, n2, m0, m1, m2);
  float a = 1.0f + 2.0f * nij;
  a = a + nij * nij;
  float d = (m0 - n0) * r0;
  d = d + (m1 - n1) * r1;
  d = d + (m2 - n2) * r2;
  float ps = a * d;
  float f = -P.A * nij * nij;
  f = f - P.B1 * nij;
  double v = (double)f - ((double)P.C * 0.25) * (double)ps;
  return (float)v;
}

// src/potential.cpp:39-43.
__device__ __forceinline__ float mcd_field(const MCDevParams &P, float elecE, float n0, float n1, float n2) {
  float d = n0 * P.eX;
  d = d + n1 * P.eY;
  d = d + n2 * P.eZ;
  float nde = elecE * d;
  return -P.elecA * (nde * nde);
}

__device__ __forceinline__ float mcd_ghrl_surface_form(const MCDevParams &P, float W, float n0, float n1, float n2,
                                                       float m0, float m1, float m2, float s0, float s1, float s2,
                                                       float ai, float aj) {
  float nij = mcd_dot(n0, n1, n2, m0, m1, m2);
  float pij = 1.5f * nij * nij - 0.5f;
  float c0 = n2 * m1 - n1 * m2;
  float c1 = n0 * m2 - n2 * m0;
  float c2 = n1 * m0 - n0 * m1;
  float cross = c0 * s0;
  cross = cross + c1 * s1;
  cross = cross + c2 * s2;
  float E1 = 1.5f * ai * ai;
  E1 = E1 + 1.5f * aj * aj;
  E1 = E1 - 1.0f;
  float t = E1 * (P.rho * pij + P.lambda);
  t = t + P.mu * (ai * aj * nij);
  t = t - 0.0f;  // the reference's integer (1/9)
  t = t + P.nu * pij;
  float sg = nij > 0 ? 1.0f : -1.0f;
  t = t + P.sigma * sg * cross;
  return W * t;
}

// Surface terms of the CPU anchoring classes (src/anchoring_homeotropic.cpp,
// src/anchoring_rp.cpp, src/anchoring_fg.cpp, src/anchoring_strong.cpp).
__device__ __forceinline__ float mcd_surface(const MCDevParams &P, int surf, float n0, float n1, float n2, float s0,
                                          float s1, float s2) {
  const float W = P.surf_W[surf];
  const float d0 = P.surf_dir[surf][0], d1 = P.surf_dir[surf][1], d2 = P.surf_dir[surf][2];
  switch (P.surf_kind[surf]) {
    case MCD_HOMEOTROPIC: {
      float nij = mcd_dot(n0, n1, n2, s0, s1, s2);
      return -W * nij * nij;
    }
    case MCD_RP:
    case MCD_STRONG:
    case MCD_STRONG_GHRL: {
      float nij = mcd_dot(n0, n1, n2, d0, d1, d2);
      return -W * nij * nij;
    }
    case MCD_FG: {
      float nij = mcd_dot(n0, n1, n2, s0, s1, s2);
      return +W * nij * nij;
    }
    case MCD_HOMEOTROPIC_GHRL: {
      float ai = mcd_dot(n0, n1, n2, s0, s1, s2);
      return mcd_ghrl_surface_form(P, W, n0, n1, n2, s0, s1, s2, s0, s1, s2, ai, 1.0f);
    }
    case MCD_RP_GHRL: {
      float ai = mcd_dot(n0, n1, n2, s0, s1, s2);
      float aj = mcd_dot(d0, d1, d2, s0, s1, s2);
      return mcd_ghrl_surface_form(P, W, n0, n1, n2, d0, d1, d2, s0, s1, s2, ai, aj);
    }
    case MCD_FG_GHRL: {
      float mod = sqrtf(fabsf(mcd_dot(n0, n1, n2, s0, s1, s2)));
      float m0 = n0 - s0 * mod, m1 = n1 - s1 * mod, m2 = n2 - s2 * mod;
      mod = sqrtf(fabsf(mcd_dot(m0, m1, m2, m0, m1, m2)));
      float dv = mod > 0 ? mod : 1.0f;
      m0 = m0 / dv;
      m1 = m1 / dv;
      m2 = m2 / dv;
      float ai = mcd_dot(n0, n1, n2, s0, s1, s2);
      float aj = mcd_dot(m0, m1, m2, s0, s1, s2);
      return mcd_ghrl_surface_form(P, W, n0, n1, n2, m0, m1, m2, s0, s1, s2, ai, aj);
    }
    default:
      return 0.0f;
  }
}

// Bond energy for neighbour b (0:+x 1:-x 2:+y 3:-y 4:+z 5:-z).
template <int POT>
__device__ __forceinline__ float mcd_bond(const MCDevParams &P, double cA, int b, float n0, float n1, float n2,
                                          float m0, float m1, float m2) {
  if (POT == MCD_LL) return mcd_bond_ll(cA, mcd_dot(n0, n1, n2, m0, m1, m2));
  const float r0 = b == 0 ? 1.0f : (b == 1 ? -1.0f : 0.0f);
  const float r1 = b == 2 ? 1.0f : (b == 3 ? -1.0f : 0.0f);
  const float r2 = b == 4 ? 1.0f : (b == 5 ? -1.0f : 0.0f);
  if (POT == MCD_GHRL) return mcd_bond_ghrl(P, n0, n1, n2, m0, m1, m2, r0, r1, r2);
  return mcd_bond_pear(P, n0, n1, n2, m0, m1, m2, r0, r1, r2);
}

// Accept rule of the reference CPU sweep, u < expf(-dE/T) (src/evolve.cpp:160,
// where exp of a float argument is glibc's expf).  A hardware exp2 decides
// every case whose margin exceeds its error bound (|x| <= 2^11 keeps the
// relative error of exp2(x*log2e) under 2^-12, and glibc's expf is within
// one ulp of exp); the rare near-ties, u == 0 and infinities are settled
// with mcx_expf, which returns glibc's expf bit for bit (mc_expf.h).
// mcx_expf's table in device global memory (one copy per translation unit).
static __device__ const uint64_t mcd_exp_table[32] = MCX_EXPF_TABLE;

// tab: the 32 entries of mcx_expf's table, wherever the caller keeps them
// (a kernel that must not wait on vector memory in this rare path copies
// them into LDS).
__device__ __forceinline__ bool mcd_accept_with(float E_new, float E_old, float T, float u, const uint64_t *tab) {
  float x = -(E_new - E_old) / T;
  float pf = __builtin_amdgcn_exp2f(x * 1.44269504088896341f);
  if (u != 0.0f && !(fabsf(u - pf) <= pf * 2.44140625e-4f)) return u < pf;
  return u < mcx_expf_with(x, tab);
}
__device__ __forceinline__ bool mcd_accept(float E_new, float E_old, float T, float u) {
  float x = -(E_new - E_old) / T;
  float pf = __builtin_amdgcn_exp2f(x * 1.44269504088896341f);
  if (u != 0.0f && !(fabsf(u - pf) <= pf * 2.44140625e-4f)) return u < pf;
  return u < mcx_expf(x);
}

// ---------------------------------------------------------------------------
// Acceptance bookkeeping shared by the sweep kernels.

// Wave64 + workgroup sum of a per-thread count.
__device__ __forceinline__ uint32_t mcd_block_sum(uint32_t v) {
  __shared__ uint32_t wsum[MCD_BLOCK / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) wsum[wid] = v;
  __syncthreads();
  uint32_t t = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < MCD_BLOCK / 64; ++w) t += wsum[w];
  return t;
}

// ang_var after a sweep with acc accepted moves of `valid` sites,
// src/evolve.cu:378-384 (= src/evolve.cpp:169-178).
__device__ __forceinline__ float mcd_adapt_ang_var(float a, uint32_t acc, int valid) {
  const double rate = 1.0 * (double)acc / (double)valid;
  if (rate < 0.5) {
    a = (float)((double)a * 0.99);
    if ((double)a < 0.01) a = (float)0.01;
  } else if (rate > 0.5) {
    a = (float)((double)a / 0.99);
    if ((double)a > 1.0) a = (float)((double)a - 0.5);
  }
  return a;
}

// Device-side ang_var adaptation by the last workgroup of a sweep's final
// pass.
//
// End of a workgroup's part of a sweep pass, called by every thread after the
// block sum (thread 0 holds block_acc).  Every pass of one sweep must use the
// same grid G.  With S = min(G, MCD_NSUB), workgroup b counts into sub-counter
// b % S: its accepted moves in the low word and, in the final pass, one
// arrival in the high word, as one 64-bit atomic.  The last arrival of a sub
// (its arrival count is known from G and S) clears it and moves the sub's
// total into the state's acc_count/blocks_done word; the last of the S subs
// to arrive there holds the sweep's count, adapts ang_var (when asked,
// src/evolve.cu:378-384), clears the word and advances the sweep index.
// Each sub sees at most ceil(G/S) atomics and the state word S.  The 64-bit
// atomics order each count with its arrival, so no fence is needed (an
// agent-scope fence in every workgroup costs an L2 write-back each); earlier
// passes of the sweep ended with their kernels, so their counts are seen.
__device__ inline void mcd_finish_sweep(const MCDevParams &P, MCDevState *st, uint32_t block_acc, bool finalize,
                                        bool adapt = true) {
  if (threadIdx.x != 0) return;
  const uint32_t G = gridDim.x, S = G < MCD_NSUB ? G : MCD_NSUB;
  const uint32_t c = blockIdx.x % S;
  unsigned long long *sub = P.ctr + (size_t)c * MCD_SUB_STRIDE;
  if (!finalize) {
    atomicAdd(sub, (unsigned long long)block_acc);
    return;
  }
  const uint32_t arrivals = G / S + (c < G % S ? 1u : 0u);
  const unsigned long long old = atomicAdd(sub, (1ull << 32) | (unsigned long long)block_acc);
  if ((uint32_t)(old >> 32) != arrivals - 1u) return;
  *sub = 0ull;
  const uint32_t part = (uint32_t)old + block_acc;
  unsigned long long *word = reinterpret_cast<unsigned long long *>(&st->acc_count);
  const unsigned long long old2 = atomicAdd(word, (1ull << 32) | (unsigned long long)part);
  if ((uint32_t)(old2 >> 32) != S - 1u) return;
  const uint32_t acc = (uint32_t)old2 + part;
  if (adapt) st->ang_var = mcd_adapt_ang_var(st->ang_var, acc, P.valid);
  st->last_acc = acc;
  *word = 0ull;  // acc_count = blocks_done = 0 for the next sweep
  st->sweep = st->sweep + 1u;
}

// End of a batched sweep for one replica, called by thread 0 of each of the
// replica's nwg workgroups: one 64-bit atomic on its acc_count/blocks_done
// word; the last arrival adapts the replica's ang_var and advances its sweep.
__device__ inline void mcd_finish_replica(const MCDevParams &P, MCDevState *st, uint32_t block_acc, uint32_t nwg) {
  unsigned long long *word = reinterpret_cast<unsigned long long *>(&st->acc_count);
  const unsigned long long old = atomicAdd(word, (1ull << 32) | (unsigned long long)block_acc);
  if ((uint32_t)(old >> 32) != nwg - 1u) return;
  const uint32_t acc = (uint32_t)old + block_acc;
  st->ang_var = mcd_adapt_ang_var(st->ang_var, acc, P.valid);
  st->last_acc = acc;
  *word = 0ull;
  st->sweep = st->sweep + 1u;
}

// Symmetric 3x3 eigenvalue closed form (Smith, CACM 4(4) 1961); largest.
__device__ inline double mcd_sym3_max(const double q[6]) {
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
  const double l2 = tr + 2.0 * pp * cos(phi + 2.0 * 3.14159265358979323846 / 3.0);
  const double l1 = 3.0 * tr - l0 - l2;
  return fmax(l0, fmax(l1, l2));
}

// Tile schedule of the persistent sweep kernels.  Tiles are the 256-item
// units one workgroup handles per iteration.  xcd == 0: plain grid stride.
// xcd == 1: the dispatcher hands workgroups to the 8 XCDs round robin, so
// workgroup b runs on XCD b % 8; each XCD gets one contiguous eighth of the
// tiles and strides through it, which keeps neighbouring rows (the halos of a
// tile) in the same XCD's L2.  Requires gridDim.x % 8 == 0 for xcd == 1.
struct MCTileRange {
  int first, end, step;
};
__device__ __forceinline__ MCTileRange mcd_tiles(int ntile, int xcd) {
  MCTileRange r;
  if (!xcd) {
    r.first = blockIdx.x;
    r.end = ntile;
    r.step = gridDim.x;
  } else {
    const int x = blockIdx.x & 7, lb = blockIdx.x >> 3;
    const int per = (ntile + 7) >> 3;
    r.first = x * per + lb;
    r.end = min(ntile, (x + 1) * per);
    r.step = gridDim.x >> 3;
  }
  return r;
}
