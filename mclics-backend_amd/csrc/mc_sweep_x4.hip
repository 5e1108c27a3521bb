// Fast checkerboard pass for all-active periodic lattices with Nx % 4 == 0
// (the bulk configurations: 256^3, the replica lattices, 128^3).
//
// Work decomposition.  A thread owns one 16-byte-aligned x segment
// [4q, 4q+4) of one row (j, k) in each SoA plane; the segment holds two sites
// of either colour.  A pass loads the segment (3 x dwordx4), the same
// segment of the four y/z neighbour rows (12 x dwordx4) and the one x
// neighbour outside the segment (3 dwords), updates its two sites of the pass
// colour and stores the segment back (3 x dwordx4) when either move is
// accepted.  Neighbouring threads read neighbouring segments, so every load
// of a wave is a contiguous 1 KiB run per plane.  Black sites written back
// during a red pass carry the values just read; no other thread writes them
// in that pass, so the in-place update is race free.
//
// Randomness.  One Philox4x32-10 block per thread and pass (counter = (index
// of the first site, sweep, stream, 0x4D44)) feeds both sites: words 0/2 give
// the angle uniforms, words 1/3 the accept uniforms (top 24 bits each) and the
// low bytes of the four words the two 16-bit axis uniforms.
//
// Accept decision.  For Lebwohl-Lasher the energy difference is first formed
// in FP32 with FMAs, sum_b c*((q.m_b)^2 - (o.m_b)^2), whose distance from the
// reference's own FP32/FP64 arithmetic is bounded by 2^-14*(1.5|A| + |field|)
// (six bonds, a few roundings each, with a 4x margin).  A hardware exp2 then
// decides every move whose uniform lies outside that error band around
// exp(-dE/T); moves inside the band (about 1e-4 of them at T ~ 1) are decided
// with the exact reference arithmetic, so every decision equals the one the
// CPU path makes.  GHRL and pear always use the exact arithmetic.
#include "mc_kernels.h"

namespace {

constexpr int kBlock = MCD_BLOCK;

__device__ __forceinline__ float sel(const float4 &v, int e) {
  return e == 0 ? v.x : (e == 1 ? v.y : (e == 2 ? v.z : v.w));
}

// 1 accept, 0 reject, -1 inside the error band (decide exactly).
__device__ __forceinline__ int fast_decide(float dE, float T, float u, float eps) {
  if (u == 0.0f) return -1;
  const float x = -dE / T;
  const float rel = eps / T + fabsf(x) * 4.76837158e-7f + 1.9073486e-6f;  // + |x| 2^-21 + 2^-19
  if (!(rel < 0.125f)) return -1;
  const float p = __builtin_amdgcn_exp2f(x * 1.44269504088896341f);
  if (fabsf(u - p) <= p * rel) return -1;
  return u < p ? 1 : 0;
}

struct Site {
  float o[3];     // current director
  float q[3];     // proposal
  float m[6][3];  // neighbours +x,-x,+y,-y,+z,-z
  float u;        // accept uniform
};

template <int POT>
__device__ __forceinline__ void exact_energies(const MCDevParams &P, double cA, float elecE, const Site &S,
                                               float &Eo, float &En) {
  double so = 0.0, sn = 0.0;
#pragma unroll
  for (int b = 0; b < 6; ++b) {
    so += (double)mcd_bond<POT>(P, cA, b, S.o[0], S.o[1], S.o[2], S.m[b][0], S.m[b][1], S.m[b][2]);
    sn += (double)mcd_bond<POT>(P, cA, b, S.q[0], S.q[1], S.q[2], S.m[b][0], S.m[b][1], S.m[b][2]);
  }
  Eo = (float)so;
  En = (float)sn;
  if (P.elecA != 0.0f) {
    Eo += mcd_field(P, elecE, S.o[0], S.o[1], S.o[2]);
    En += mcd_field(P, elecE, S.q[0], S.q[1], S.q[2]);
  }
}

template <int POT>
__device__ __forceinline__ bool decide(const MCDevParams &P, double cA, float cAf, float eps, float T, float elecE,
                                       const Site &S) {
  if (POT == MCD_LL) {
    float acc = 0.0f;
#pragma unroll
    for (int b = 0; b < 6; ++b) {
      const float a = fmaf(S.o[2], S.m[b][2], fmaf(S.o[1], S.m[b][1], S.o[0] * S.m[b][0]));
      const float c = fmaf(S.q[2], S.m[b][2], fmaf(S.q[1], S.m[b][1], S.q[0] * S.m[b][0]));
      acc = fmaf(c - a, c + a, acc);
    }
    float dE = cAf * acc;
    if (P.elecA != 0.0f)
      dE += mcd_field(P, elecE, S.q[0], S.q[1], S.q[2]) - mcd_field(P, elecE, S.o[0], S.o[1], S.o[2]);
    const int d = fast_decide(dE, T, S.u, eps);
    if (d >= 0) return d == 1;
  }
  float Eo, En;
  exact_energies<POT>(P, cA, elecE, S, Eo, En);
  return mcd_accept(En, Eo, T, S.u);
}

// Proposal of src/evolve.cu:143-163 with a 16-bit axis uniform.
__device__ __forceinline__ void propose(const float o[3], float av, uint32_t w_angle, uint32_t axis16, float q[3]) {
  const float va = (2.0f * mcd_u01(w_angle) - 1.0f) * av;
  const float h = 0.5f * va;  // revolutions: sin(pi*va) = sin(2*pi*h)
  const float sn = __builtin_amdgcn_sinf(h), cs = __builtin_amdgcn_cosf(h);
  float q0, q1, q2;
  if (axis16 < 21824u) {  // u < 0.333
    q0 = o[0];
    q1 = o[1] * cs + o[2] * sn;
    q2 = o[2] * cs - o[1] * sn;
  } else if (axis16 < 43647u) {  // u < 0.666
    q0 = o[0] * cs - o[2] * sn;
    q1 = o[1];
    q2 = o[2] * cs + o[0] * sn;
  } else {
    q0 = o[0] * cs + o[1] * sn;
    q1 = o[1] * cs - o[0] * sn;
    q2 = o[2];
  }
  const float inv = __builtin_amdgcn_rsqf(q0 * q0 + q1 * q1 + q2 * q2);
  q[0] = q0 * inv;
  q[1] = q1 * inv;
  q[2] = q2 * inv;
}

// One thread's segment update; PAR is the x parity of the first site of the
// pass colour in this row (sites at 4q+PAR and 4q+PAR+2).
template <int POT, bool RECORD, int PAR>
__device__ __forceinline__ uint32_t segment(const MCDevParams &P, float *__restrict__ nx, float *__restrict__ ny,
                                            float *__restrict__ nz, int q, int j, int k, uint32_t sweep, float T,
                                            float elecE, float av, double cA, float cAf, float eps,
                                            const MCRecord &rec) {
  const int X = P.nx, Y = P.ny, Z = P.nz;
  const int nseg = X >> 2;
  const int row = X * (j + Y * k);
  const int jp = j + 1 == Y ? 0 : j + 1, jm = j == 0 ? Y - 1 : j - 1;
  const int kp = k + 1 == Z ? 0 : k + 1, km = k == 0 ? Z - 1 : k - 1;
  const int x0 = 4 * q;
  const int r_jp = X * (jp + Y * k) + x0, r_jm = X * (jm + Y * k) + x0;
  const int r_kp = X * (j + Y * kp) + x0, r_km = X * (j + Y * km) + x0;
  const int own = row + x0;
  const int out = PAR == 0 ? row + (q == 0 ? X - 1 : x0 - 1) : row + (q == nseg - 1 ? 0 : x0 + 4);
  float *pl[3] = {nx, ny, nz};
  float4 c[3], yp[3], ym[3], zp[3], zm[3];
  float ov[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    c[d] = *reinterpret_cast<const float4 *>(pl[d] + own);
    yp[d] = *reinterpret_cast<const float4 *>(pl[d] + r_jp);
    ym[d] = *reinterpret_cast<const float4 *>(pl[d] + r_jm);
    zp[d] = *reinterpret_cast<const float4 *>(pl[d] + r_kp);
    zm[d] = *reinterpret_cast<const float4 *>(pl[d] + r_km);
    ov[d] = pl[d][out];
  }
  const int sA = own + PAR;  // linear index of the first site
  const MCU4 r = mcd_philox(MCU4{(uint32_t)sA, sweep, P.stream, 0x4D44u}, P.seed_lo, P.seed_hi);
  Site S[2];
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const int e = PAR + 2 * t;  // element of this site in the segment
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      S[t].o[d] = sel(c[d], e);
      S[t].m[0][d] = (e == 3) ? ov[d] : sel(c[d], e + 1);  // +x
      S[t].m[1][d] = (e == 0) ? ov[d] : sel(c[d], e - 1);  // -x
      S[t].m[2][d] = sel(yp[d], e);
      S[t].m[3][d] = sel(ym[d], e);
      S[t].m[4][d] = sel(zp[d], e);
      S[t].m[5][d] = sel(zm[d], e);
    }
  }
  const uint32_t ax0 = ((r.x & 0xFFu) << 8) | (r.y & 0xFFu);
  const uint32_t ax1 = ((r.z & 0xFFu) << 8) | (r.w & 0xFFu);
  propose(S[0].o, av, r.x, ax0, S[0].q);
  propose(S[1].o, av, r.z, ax1, S[1].q);
  S[0].u = mcd_u01(r.y);
  S[1].u = mcd_u01(r.w);
  bool a[2];
#pragma unroll
  for (int t = 0; t < 2; ++t) a[t] = decide<POT>(P, cA, cAf, eps, T, elecE, S[t]);
  if (a[0] || a[1]) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      float4 v = c[d];
      if (a[0]) {
        if (PAR == 0) v.x = S[0].q[d]; else v.y = S[0].q[d];
      }
      if (a[1]) {
        if (PAR == 0) v.z = S[1].q[d]; else v.w = S[1].q[d];
      }
      *reinterpret_cast<float4 *>(pl[d] + own) = v;
    }
  }
  if (RECORD) {
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const int s = sA + 2 * t;
      float Eo, En;
      exact_energies<POT>(P, cA, elecE, S[t], Eo, En);
      rec.u0[s] = mcd_u01(t ? r.z : r.x);
      rec.u1[s] = (float)(t ? ax1 : ax0) * (1.0f / 65536.0f);
      rec.u2[s] = S[t].u;
      rec.px[s] = S[t].q[0];
      rec.py[s] = S[t].q[1];
      rec.pz[s] = S[t].q[2];
      rec.acc[s] = a[t] ? 1 : 0;
      rec.eo[s] = Eo;
      rec.en[s] = En;
    }
  }
  return (a[0] ? 1u : 0u) + (a[1] ? 1u : 0u);
}

template <int POT, bool RECORD>
__global__ __launch_bounds__(kBlock) void k_sweep_x4(MCDevParams P, float *__restrict__ nx, float *__restrict__ ny,
                                                     float *__restrict__ nz, MCDevState *st, int colour, int finalize,
                                                     int xcd, MCRecord rec) {
  const float T = st->T, elecE = st->elecE, av = st->ang_var;
  const uint32_t sweep = st->sweep;
  const double cA = (double)(-P.A) * 1.5;
  const float cAf = (float)cA;
  const float e2 = P.eX * P.eX + P.eY * P.eY + P.eZ * P.eZ;
  const float eps = 6.1035156e-5f * (fabsf(cAf) + fabsf(P.elecA) * elecE * elecE * e2);  // 2^-14 * (...)
  const int nseg = P.nx >> 2;
  const int total = nseg * P.ny * P.nz;
  uint32_t acc = 0;
  const MCTileRange tr = mcd_tiles((total + kBlock - 1) / kBlock, xcd);
  for (int tile = tr.first; tile < tr.end; tile += tr.step) {
    const int t = tile * kBlock + threadIdx.x;
    if (t >= total) continue;
    const int q = t % nseg;
    const int rowi = t / nseg;
    const int j = rowi % P.ny, k = rowi / P.ny;
    if (((j + k + colour) & 1) == 0)
      acc += segment<POT, RECORD, 0>(P, nx, ny, nz, q, j, k, sweep, T, elecE, av, cA, cAf, eps, rec);
    else
      acc += segment<POT, RECORD, 1>(P, nx, ny, nz, q, j, k, sweep, T, elecE, av, cA, cAf, eps, rec);
  }
  const uint32_t bsum = mcd_block_sum(acc);
  mcd_finish_sweep(P, st, bsum, finalize != 0);
}

template <int POT>
void launch_t(const MCLaunch &L, int colour, int grid, int finalize, int xcd, const MCRecord *rec) {
  if (rec)
    k_sweep_x4<POT, true><<<grid, kBlock, 0, L.stream>>>(L.P, L.nx, L.ny, L.nz, L.st, colour, finalize, xcd, *rec);
  else
    k_sweep_x4<POT, false><<<grid, kBlock, 0, L.stream>>>(L.P, L.nx, L.ny, L.nz, L.st, colour, finalize, xcd,
                                                          MCRecord{});
}

}  // namespace

bool mc_x4_eligible(const MCLaunch &L) {
  return L.pt == nullptr && L.P.n_surf == 0 && (L.P.nx % 4) == 0 && L.P.per[0] && L.P.per[1] && L.P.per[2];
}

int mc_x4_slots(const MCDevParams &P) { return (P.nx / 4) * P.ny * P.nz; }

void mc_launch_sweep_x4(const MCLaunch &L, int colour, int grid, int finalize, int xcd, const MCRecord *rec) {
  switch (L.P.potential) {
    case MCD_GHRL:
      launch_t<MCD_GHRL>(L, colour, grid, finalize, xcd, rec);
      break;
    case MCD_PEAR:
      launch_t<MCD_PEAR>(L, colour, grid, finalize, xcd, rec);
      break;
    default:
      launch_t<MCD_LL>(L, colour, grid, finalize, xcd, rec);
      break;
  }
}

// ---------------------------------------------------------------------------
// Batched replicas: R independent lattices (uneven sizes, each all active,
// periodic, Nx % 4 == 0) in one set of SoA planes.  A pass is one launch over
// the tiles of every replica (a tile = the 256 segments one workgroup updates
// in one iteration; tiles never straddle replicas, so each workgroup reduces
// its acceptance count into exactly one replica counter per tile).  The
// replica of a tile comes from a tile table; each replica keeps its own
// parameters, temperature, proposal width and generator stream, and evolves
// exactly as the same lattice would alone in the segment kernel.

namespace {

template <int POT>
__global__ __launch_bounds__(kBlock) void k_batch_x4(const MCDevParams *__restrict__ params, MCDevState *states,
                                                     const int *__restrict__ tile_rep, const int *__restrict__ rep_tile0,
                                                     const long long *__restrict__ rep_off, int ntiles, int nrep,
                                                     float *__restrict__ nx, float *__restrict__ ny,
                                                     float *__restrict__ nz, int colour, int finalize, int xcd,
                                                     uint32_t *blocks_done) {
  const MCTileRange tr = mcd_tiles(ntiles, xcd);
  for (int tile = tr.first; tile < tr.end; tile += tr.step) {
    const int r = tile_rep[tile];
    const MCDevParams P = params[r];
    MCDevState *st = states + r;
    const float T = st->T, elecE = st->elecE, av = st->ang_var;
    const uint32_t sweep = st->sweep;
    const double cA = (double)(-P.A) * 1.5;
    const float cAf = (float)cA;
    const float e2 = P.eX * P.eX + P.eY * P.eY + P.eZ * P.eZ;
    const float eps = 6.1035156e-5f * (fabsf(cAf) + fabsf(P.elecA) * elecE * elecE * e2);
    const long long off = rep_off[r];
    const int nseg = P.nx >> 2;
    const int total = nseg * P.ny * P.nz;
    const int t = (tile - rep_tile0[r]) * kBlock + threadIdx.x;
    uint32_t acc = 0;
    if (t < total) {
      const int q = t % nseg;
      const int rowi = t / nseg;
      const int j = rowi % P.ny, k = rowi / P.ny;
      if (((j + k + colour) & 1) == 0)
        acc = segment<POT, false, 0>(P, nx + off, ny + off, nz + off, q, j, k, sweep, T, elecE, av, cA, cAf, eps,
                                     MCRecord{});
      else
        acc = segment<POT, false, 1>(P, nx + off, ny + off, nz + off, q, j, k, sweep, T, elecE, av, cA, cAf, eps,
                                     MCRecord{});
    }
    const uint32_t bsum = mcd_block_sum(acc);
    if (threadIdx.x == 0) atomicAdd(&st->acc_count, bsum);
    __syncthreads();  // wsum of mcd_block_sum is reused by the next tile
  }
  if (!finalize) return;
  __shared__ bool last;
  if (threadIdx.x == 0) {
    __threadfence();
    last = atomicAdd(blocks_done, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  for (int r = threadIdx.x; r < nrep; r += kBlock) {
    MCDevState *st = states + r;
    const uint32_t acc = atomicAdd(&st->acc_count, 0u);
    float a = st->ang_var;
    const double rate = 1.0 * (double)acc / (double)params[r].valid;
    if (rate < 0.5) {
      a = (float)((double)a * 0.99);
      if ((double)a < 0.01) a = (float)0.01;
    } else if (rate > 0.5) {
      a = (float)((double)a / 0.99);
      if ((double)a > 1.0) a = (float)((double)a - 0.5);
    }
    st->ang_var = a;
    st->last_acc = acc;
    st->acc_count = 0u;
    st->sweep = st->sweep + 1u;
  }
  if (threadIdx.x == 0) *blocks_done = 0u;
}

}  // namespace

void mc_launch_batch_x4(const MCBatchLaunch &B, int colour, int grid, int finalize) {
  switch (B.potential) {
    case MCD_GHRL:
      k_batch_x4<MCD_GHRL><<<grid, kBlock, 0, B.stream>>>(B.params, B.states, B.tile_rep, B.rep_tile0, B.rep_off,
                                                          B.ntiles, B.nrep, B.nx, B.ny, B.nz, colour, finalize,
                                                          B.xcd, B.blocks_done);
      break;
    case MCD_PEAR:
      k_batch_x4<MCD_PEAR><<<grid, kBlock, 0, B.stream>>>(B.params, B.states, B.tile_rep, B.rep_tile0, B.rep_off,
                                                          B.ntiles, B.nrep, B.nx, B.ny, B.nz, colour, finalize,
                                                          B.xcd, B.blocks_done);
      break;
    default:
      k_batch_x4<MCD_LL><<<grid, kBlock, 0, B.stream>>>(B.params, B.states, B.tile_rep, B.rep_tile0, B.rep_off,
                                                        B.ntiles, B.nrep, B.nx, B.ny, B.nz, colour, finalize,
                                                        B.xcd, B.blocks_done);
      break;
  }
}
