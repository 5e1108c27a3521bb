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
// in that pass, so the in-place update is race free.  Proposal, randomness
// and accept rule are those of mc_site_update.h.
#include "mc_kernels.h"
#include "mc_site_update.h"

namespace {

constexpr int kBlock = MCD_BLOCK;
using mcsu::sel;

// Row of segment-row index rowi.  With fewer than 64 segments per row a wave
// spans several rows, and rows of opposite (j + k) parity take the two
// different PAR branches; so the rows with (j + k) even come first, then the
// odd ones, and every wave stays on one parity (when (Ny Nz / 2) * Nx/4 is a
// multiple of 64).  Ny is even (periodic axes are).  The generator is keyed
// by site, so the order changes nothing but the speed.
__device__ __forceinline__ void x4_row(int rowi, int nseg, int Y, int Z, int &j, int &k) {
  if (nseg >= 64) {
    j = rowi % Y;
    k = rowi / Y;
    return;
  }
  const int h = (Y * Z) >> 1, hy = Y >> 1;
  const int odd = rowi >= h ? 1 : 0;
  const int e = rowi - odd * h;
  k = e / hy;
  j = 2 * (e - k * hy) + ((k + odd) & 1);
}

// The loads of one segment (slot t) for a pass of the given colour: the
// segment, the pass-colour sites of its four y/z neighbour segments and the
// x neighbour outside it.  PAR is the x parity of the first site of the pass
// colour in the row (sites at 4q+PAR and 4q+PAR+2).  Only sites of the colour
// the pass does not update are read from other segments, and the segment
// itself belongs to this thread alone, so loads issued ahead of earlier
// updates see the pass's input.
struct SegIn {
  float4 c[3];
  mcsu::PairNbrs nb;
  int q, j, k, par;
  bool valid;
};

template <bool GEO>
__device__ __forceinline__ void seg_load(const MCDevParams &P, float *const pl[3], int t, int total, int colour,
                                         SegIn &s) {
  s.valid = t < total;
  if (!s.valid) return;
  const int X = P.nx, Y = P.ny, Z = P.nz;
  const int nseg = X >> 2;
  s.q = t % nseg;
  x4_row(t / nseg, nseg, Y, Z, s.j, s.k);
  const int q = s.q, j = s.j, k = s.k;
  s.par = (j + k + colour) & 1;
  const int row = X * (j + Y * k);
  const int jp = j + 1 == Y ? 0 : j + 1, jm = j == 0 ? Y - 1 : j - 1;
  const int kp = k + 1 == Z ? 0 : k + 1, km = k == 0 ? Z - 1 : k - 1;
  const int x0 = 4 * q;
  const int r_jp = X * (jp + Y * k) + x0, r_jm = X * (jm + Y * k) + x0;
  const int r_kp = X * (j + Y * kp) + x0, r_km = X * (j + Y * km) + x0;
  const int out = s.par == 0 ? row + (q == 0 ? X - 1 : x0 - 1) : row + (q == nseg - 1 ? 0 : x0 + 4);
  // free z faces: the missing neighbour is a zero vector
  const bool zp_on = !GEO || !P.zfree || k + 1 < Z, zm_on = !GEO || !P.zfree || k > 0;
  const float4 zero4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    s.c[d] = *reinterpret_cast<const float4 *>(pl[d] + row + x0);
    const float4 yp = *reinterpret_cast<const float4 *>(pl[d] + r_jp);
    const float4 ym = *reinterpret_cast<const float4 *>(pl[d] + r_jm);
    const float4 zp = zp_on ? *reinterpret_cast<const float4 *>(pl[d] + r_kp) : zero4;
    const float4 zm = zm_on ? *reinterpret_cast<const float4 *>(pl[d] + r_km) : zero4;
    if (s.par == 0) {
      s.nb.yp[d] = make_float2(yp.x, yp.z);
      s.nb.ym[d] = make_float2(ym.x, ym.z);
      s.nb.zp[d] = make_float2(zp.x, zp.z);
      s.nb.zm[d] = make_float2(zm.x, zm.z);
    } else {
      s.nb.yp[d] = make_float2(yp.y, yp.w);
      s.nb.ym[d] = make_float2(ym.y, ym.w);
      s.nb.zp[d] = make_float2(zp.y, zp.w);
      s.nb.zm[d] = make_float2(zm.y, zm.w);
    }
    s.nb.ov[d] = pl[d][out];
  }
}

// One thread's segment update from its loads.
template <int POT, bool RECORD, int PAR, bool GEO>
__device__ __forceinline__ uint32_t seg_update(const MCDevParams &P, const mcsu::PassConst &kc, float *const pl[3],
                                               const SegIn &s, const MCRecord &rec) {
  const int X = P.nx, Y = P.ny, Z = P.nz;
  const int k = s.k;
  const int own = X * (s.j + Y * k) + 4 * s.q;
  mcsu::Where w = mcsu::kBulk;
  const bool zp_on = !GEO || !P.zfree || k + 1 < Z, zm_on = !GEO || !P.zfree || k > 0;
  if (!zp_on) w.on &= ~(1u << 4);
  if (!zm_on) w.on &= ~(1u << 5);
  if (GEO && P.zsurf) w.surf = k == 0 ? 0 : (k == Z - 1 ? 1 : -1);
  float4 c[3] = {s.c[0], s.c[1], s.c[2]};
  const int sA = own + PAR;  // linear index of the first site
  mcsu::Site S[2];
  bool a[2];
  MCU4 r;
  mcsu::pair_update<POT, PAR, GEO>(P, kc, c, s.nb, (uint32_t)sA, w, S, a, r);
  if (a[0] || a[1]) {
    mcsu::pair_apply<PAR>(c, S, a);
#pragma unroll
    for (int d = 0; d < 3; ++d) *reinterpret_cast<float4 *>(pl[d] + own) = c[d];
  }
  if (RECORD) mcsu::pair_record<POT, GEO>(P, kc, rec, sA, w, S, a, r);
  return (a[0] ? 1u : 0u) + (a[1] ? 1u : 0u);
}

template <int POT, bool RECORD, bool GEO>
__device__ __forceinline__ uint32_t seg_update_any(const MCDevParams &P, const mcsu::PassConst &kc,
                                                   float *const pl[3], const SegIn &s, const MCRecord &rec) {
  if (s.par == 0) return seg_update<POT, RECORD, 0, GEO>(P, kc, pl, s, rec);
  return seg_update<POT, RECORD, 1, GEO>(P, kc, pl, s, rec);
}

// PIPE: the workgroup has several tiles; the next tile's loads are issued
// before the current segment is updated (one more segment of registers).
template <int POT, bool RECORD, bool GEO, bool PIPE>
__global__ __launch_bounds__(kBlock) void k_sweep_x4(MCDevParams P, float *__restrict__ nx, float *__restrict__ ny,
                                                     float *__restrict__ nz, MCDevState *st, int colour, int finalize,
                                                     int xcd, MCRecord rec) {
  const mcsu::PassConst kc = mcsu::pass_const(P, st);
  const int nseg = P.nx >> 2;
  const int total = nseg * P.ny * P.nz;
  float *const pl[3] = {nx, ny, nz};
  uint32_t acc = 0;
  const MCTileRange tr = mcd_tiles((total + kBlock - 1) / kBlock, xcd);
  if (PIPE) {
    SegIn cur, nxt;
    cur.valid = false;
    if (tr.first < tr.end) seg_load<GEO>(P, pl, tr.first * kBlock + (int)threadIdx.x, total, colour, cur);
    for (int tile = tr.first; tile < tr.end; tile += tr.step) {
      nxt.valid = false;
      if (tile + tr.step < tr.end)
        seg_load<GEO>(P, pl, (tile + tr.step) * kBlock + (int)threadIdx.x, total, colour, nxt);
      if (cur.valid) acc += seg_update_any<POT, RECORD, GEO>(P, kc, pl, cur, rec);
      cur = nxt;
    }
  } else {
    for (int tile = tr.first; tile < tr.end; tile += tr.step) {
      SegIn sg;
      seg_load<GEO>(P, pl, tile * kBlock + (int)threadIdx.x, total, colour, sg);
      if (sg.valid) acc += seg_update_any<POT, RECORD, GEO>(P, kc, pl, sg, rec);
    }
  }
  const uint32_t bsum = mcd_block_sum(acc);
  mcd_finish_sweep(P, st, bsum, finalize != 0);
}

// Measurement of an all-active periodic lattice with Nx % 4 == 0: mean site
// energy (each site's energy with the reference arithmetic, src/geometry.cpp:
// 19-72, summed in FP64) and the FP64 Q tensor, one 16-byte segment per thread
// over the same XCD-contiguous tile order as the sweeps, so the y/z
// neighbour rows come from L2.  Writes FP64 partials per workgroup
// {E, Qxx, Qxy, Qxz, Qyy, Qyz, Qzz, n}; the one-block finisher of the general
// measurement reduces them in workgroup order (deterministic).
// Adds the four sites of segment q of row (j, k) to a[8] = {E, 3xx-1, 3xy,
// 3xz, 3yy-1, 3yz, 3zz-1, n}: each site energy with the reference arithmetic
// (FP32 bonds summed in FP64, then rounded to FP32), Q terms in FP64.
template <int POT, bool GEO>
__device__ __forceinline__ void measure_segment(const MCDevParams &P, double cA, float elecE,
                                                const float *const pl[3], int q, int j, int k, double a[8]) {
  const int X = P.nx, Y = P.ny, Z = P.nz;
  const int nseg = X >> 2;
  const int jp = j + 1 == Y ? 0 : j + 1, jm = j == 0 ? Y - 1 : j - 1;
  const int kp = k + 1 == Z ? 0 : k + 1, km = k == 0 ? Z - 1 : k - 1;
  const int row = X * (j + Y * k), x0 = 4 * q;
  const int r_jp = X * (jp + Y * k) + x0, r_jm = X * (jm + Y * k) + x0;
  const int r_kp = X * (j + Y * kp) + x0, r_km = X * (j + Y * km) + x0;
  const int xl = row + (q == 0 ? X - 1 : x0 - 1), xr = row + (q == nseg - 1 ? 0 : x0 + 4);
  const bool zp_on = !GEO || !P.zfree || k + 1 < Z, zm_on = !GEO || !P.zfree || k > 0;
  const int surf = GEO && P.zsurf ? (k == 0 ? 0 : (k == Z - 1 ? 1 : -1)) : -1;
  const float4 zero4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  float4 c[3], yp[3], ym[3], zp[3], zm[3];
  float ol[3], orr[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    c[d] = *reinterpret_cast<const float4 *>(pl[d] + row + x0);
    yp[d] = *reinterpret_cast<const float4 *>(pl[d] + r_jp);
    ym[d] = *reinterpret_cast<const float4 *>(pl[d] + r_jm);
    zp[d] = zp_on ? *reinterpret_cast<const float4 *>(pl[d] + r_kp) : zero4;
    zm[d] = zm_on ? *reinterpret_cast<const float4 *>(pl[d] + r_km) : zero4;
    ol[d] = pl[d][xl];
    orr[d] = pl[d][xr];
  }
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    float o[3], m[6][3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      o[d] = sel(c[d], e);
      m[0][d] = e == 3 ? orr[d] : sel(c[d], e + 1);  // +x
      m[1][d] = e == 0 ? ol[d] : sel(c[d], e - 1);   // -x
      m[2][d] = sel(yp[d], e);
      m[3][d] = sel(ym[d], e);
      m[4][d] = sel(zp[d], e);
      m[5][d] = sel(zm[d], e);
    }
    double so = 0.0;
#pragma unroll
    for (int b = 0; b < 6; ++b) {
      if ((b == 4 && !zp_on) || (b == 5 && !zm_on)) continue;
      so += (double)mcd_bond<POT>(P, cA, b, o[0], o[1], o[2], m[b][0], m[b][1], m[b][2]);
    }
    float E = (float)so;
    if (GEO && surf >= 0) E += mcd_surface(P, surf, o[0], o[1], o[2], 0.0f, 0.0f, surf ? 1.0f : -1.0f);
    if (P.elecA != 0.0f) E += mcd_field(P, elecE, o[0], o[1], o[2]);
    const double x = o[0], y = o[1], z = o[2];
    a[0] += (double)E;
    a[1] += 3.0 * x * x - 1.0;
    a[2] += 3.0 * x * y;
    a[3] += 3.0 * x * z;
    a[4] += 3.0 * y * y - 1.0;
    a[5] += 3.0 * y * z;
    a[6] += 3.0 * z * z - 1.0;
    a[7] += 1.0;
  }
}

// Sum over the workgroup of a[8]; threads 0..7 get component threadIdx.x.
__device__ __forceinline__ double block_sum8(const double a[8]) {
  __shared__ double red[8][kBlock / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    double v = a[c];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if (lane == 0) red[c][wid] = v;
  }
  __syncthreads();
  double v = 0;
  if (threadIdx.x < 8)
    for (int w = 0; w < kBlock / 64; ++w) v += red[threadIdx.x][w];
  __syncthreads();  // red is reused by the next call
  return v;
}

template <int POT, bool GEO>
__global__ __launch_bounds__(kBlock) void k_measure_x4(MCDevParams P, const float *__restrict__ nx,
                                                       const float *__restrict__ ny, const float *__restrict__ nz,
                                                       const MCDevState *st, int xcd,
                                                       double *__restrict__ partials) {
  const float elecE = st->elecE;
  const double cA = (double)(-P.A) * 1.5;
  const int nseg = P.nx >> 2;
  const int total = nseg * P.ny * P.nz;
  const float *const pl[3] = {nx, ny, nz};
  double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const MCTileRange tr = mcd_tiles((total + kBlock - 1) / kBlock, xcd);
  for (int tile = tr.first; tile < tr.end; tile += tr.step) {
    const int t = tile * kBlock + threadIdx.x;
    if (t >= total) continue;
    const int q = t % nseg, rowi = t / nseg;
    measure_segment<POT, GEO>(P, cA, elecE, pl, q, rowi % P.ny, rowi / P.ny, a);
  }
  const double v = block_sum8(a);
  if (threadIdx.x < 8) partials[(size_t)blockIdx.x * 8 + threadIdx.x] = v;
}

// Batched replicas: FP64 partial sums {E, Qxx..Qzz, n} per tile (256
// segments of one replica), reduced per replica in tile order by
// k_batch_measure_final (mc_kernels.hip).
template <int POT>
__global__ __launch_bounds__(kBlock) void k_batch_measure_x4(const MCDevParams *__restrict__ params,
                                                             const MCDevState *__restrict__ states,
                                                             const int *__restrict__ tile_rep,
                                                             const int *__restrict__ rep_tile0,
                                                             const long long *__restrict__ rep_off, int ntiles,
                                                             const float *__restrict__ nx,
                                                             const float *__restrict__ ny,
                                                             const float *__restrict__ nz,
                                                             double *__restrict__ partials) {
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int r = tile_rep[tile];
    const MCDevParams P = params[r];
    const float elecE = states[r].elecE;
    const double cA = (double)(-P.A) * 1.5;
    const long long off = rep_off[r];
    const float *const pl[3] = {nx + off, ny + off, nz + off};
    const int nseg = P.nx >> 2;
    const int t = (tile - rep_tile0[r]) * kBlock + threadIdx.x;
    double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (t < nseg * P.ny * P.nz) {
      const int q = t % nseg, rowi = t / nseg;
      measure_segment<POT, false>(P, cA, elecE, pl, q, rowi % P.ny, rowi / P.ny, a);
    }
    const double v = block_sum8(a);
    if (threadIdx.x < 8) partials[(size_t)tile * 8 + threadIdx.x] = v;
  }
}

template <int POT, bool GEO>
void launch_g(const MCLaunch &L, int colour, int grid, int finalize, int xcd, const MCRecord *rec) {
  // the pipelined loop when some workgroup has two or more tiles
  const bool pipe = grid < (mc_x4_slots(L.P) + kBlock - 1) / kBlock;
  if (rec)
    k_sweep_x4<POT, true, GEO, false><<<grid, kBlock, 0, L.stream>>>(L.P, L.nx, L.ny, L.nz, L.st, colour, finalize,
                                                                      xcd, *rec);
  else if (pipe)
    k_sweep_x4<POT, false, GEO, true><<<grid, kBlock, 0, L.stream>>>(L.P, L.nx, L.ny, L.nz, L.st, colour, finalize,
                                                                      xcd, MCRecord{});
  else
    k_sweep_x4<POT, false, GEO, false><<<grid, kBlock, 0, L.stream>>>(L.P, L.nx, L.ny, L.nz, L.st, colour, finalize,
                                                                       xcd, MCRecord{});
}

template <int POT>
void launch_t(const MCLaunch &L, int colour, int grid, int finalize, int xcd, const MCRecord *rec) {
  if (L.P.zfree)
    launch_g<POT, true>(L, colour, grid, finalize, xcd, rec);
  else
    launch_g<POT, false>(L, colour, grid, finalize, xcd, rec);
}

template <int POT>
void launch_measure_t(const MCLaunch &L, int grid, int xcd, double *partials) {
  if (L.P.zfree)
    k_measure_x4<POT, true><<<grid, kBlock, 0, L.stream>>>(L.P, L.nx, L.ny, L.nz, L.st, xcd, partials);
  else
    k_measure_x4<POT, false><<<grid, kBlock, 0, L.stream>>>(L.P, L.nx, L.ny, L.nz, L.st, xcd, partials);
}

}  // namespace

bool mc_x4_eligible(const MCLaunch &L) {
  if ((L.P.nx % 4) != 0 || !L.P.per[0] || !L.P.per[1]) return false;
  if (L.P.zfree) return true;  // slab geometry, verified against pt and ns at create
  return L.pt == nullptr && L.P.n_surf == 0 && L.P.per[2];
}

int mc_x4_slots(const MCDevParams &P) { return (P.nx / 4) * P.ny * P.nz; }

void mc_launch_batch_measure_tiles(const MCBatchLaunch &B, int grid, double *tile_partials) {
  switch (B.potential) {
    case MCD_GHRL:
      k_batch_measure_x4<MCD_GHRL><<<grid, kBlock, 0, B.stream>>>(B.params, B.states, B.tile_rep, B.rep_tile0,
                                                                  B.rep_off, B.ntiles, B.nx, B.ny, B.nz, tile_partials);
      break;
    case MCD_PEAR:
      k_batch_measure_x4<MCD_PEAR><<<grid, kBlock, 0, B.stream>>>(B.params, B.states, B.tile_rep, B.rep_tile0,
                                                                  B.rep_off, B.ntiles, B.nx, B.ny, B.nz, tile_partials);
      break;
    default:
      k_batch_measure_x4<MCD_LL><<<grid, kBlock, 0, B.stream>>>(B.params, B.states, B.tile_rep, B.rep_tile0,
                                                                B.rep_off, B.ntiles, B.nx, B.ny, B.nz, tile_partials);
      break;
  }
}

void mc_launch_measure_x4(const MCLaunch &L, int grid, int xcd, double *partials, double *out, double *series) {
  switch (L.P.potential) {
    case MCD_GHRL:
      launch_measure_t<MCD_GHRL>(L, grid, xcd, partials);
      break;
    case MCD_PEAR:
      launch_measure_t<MCD_PEAR>(L, grid, xcd, partials);
      break;
    default:
      launch_measure_t<MCD_LL>(L, grid, xcd, partials);
      break;
  }
  mc_launch_measure_final(L.stream, partials, grid, out, series);
}

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
                                                     float *__restrict__ nz, int colour, int xcd) {
  const MCTileRange tr = mcd_tiles(ntiles, xcd);
  for (int tile = tr.first; tile < tr.end; tile += tr.step) {
    const int r = tile_rep[tile];
    const MCDevParams P = params[r];
    MCDevState *st = states + r;
    const mcsu::PassConst kc = mcsu::pass_const(P, st);
    const long long off = rep_off[r];
    const int nseg = P.nx >> 2;
    const int total = nseg * P.ny * P.nz;
    const int t = (tile - rep_tile0[r]) * kBlock + threadIdx.x;
    uint32_t acc = 0;
    float *const pl[3] = {nx + off, ny + off, nz + off};
    SegIn sg;
    seg_load<false>(P, pl, t, total, colour, sg);
    if (sg.valid) acc = seg_update_any<POT, false, false>(P, kc, pl, sg, MCRecord{});
    const uint32_t bsum = mcd_block_sum(acc);
    if (threadIdx.x == 0) atomicAdd(&st->acc_count, bsum);
    __syncthreads();  // wsum of mcd_block_sum is reused by the next tile
  }
}

// End of a batched sweep: per-replica ang_var adaptation from the counts the
// passes accumulated (a separate one-block launch, so the counts are complete
// and visible without per-workgroup fences), src/evolve.cu:378-384.
__global__ __launch_bounds__(kBlock) void k_batch_finish(const MCDevParams *__restrict__ params, MCDevState *states,
                                                         int nrep) {
  for (int r = threadIdx.x; r < nrep; r += kBlock) {
    MCDevState *st = states + r;
    const uint32_t acc = st->acc_count;
    st->ang_var = mcd_adapt_ang_var(st->ang_var, acc, params[r].valid);
    st->last_acc = acc;
    st->acc_count = 0u;
    st->sweep = st->sweep + 1u;
  }
}

}  // namespace

void mc_launch_batch_x4(const MCBatchLaunch &B, int colour, int grid, int finalize) {
  switch (B.potential) {
    case MCD_GHRL:
      k_batch_x4<MCD_GHRL><<<grid, kBlock, 0, B.stream>>>(B.params, B.states, B.tile_rep, B.rep_tile0, B.rep_off,
                                                          B.ntiles, B.nrep, B.nx, B.ny, B.nz, colour, B.xcd);
      break;
    case MCD_PEAR:
      k_batch_x4<MCD_PEAR><<<grid, kBlock, 0, B.stream>>>(B.params, B.states, B.tile_rep, B.rep_tile0, B.rep_off,
                                                          B.ntiles, B.nrep, B.nx, B.ny, B.nz, colour, B.xcd);
      break;
    default:
      k_batch_x4<MCD_LL><<<grid, kBlock, 0, B.stream>>>(B.params, B.states, B.tile_rep, B.rep_tile0, B.rep_off,
                                                        B.ntiles, B.nrep, B.nx, B.ny, B.nz, colour, B.xcd);
      break;
  }
  if (finalize) mc_launch_batch_finish(B);
}

void mc_launch_batch_finish(const MCBatchLaunch &B) {
  k_batch_finish<<<1, kBlock, 0, B.stream>>>(B.params, B.states, B.nrep);
}
