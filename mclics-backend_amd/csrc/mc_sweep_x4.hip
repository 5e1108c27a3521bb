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

// One thread's segment update; PAR is the x parity of the first site of the
// pass colour in this row (sites at 4q+PAR and 4q+PAR+2).
template <int POT, bool RECORD, int PAR>
__device__ __forceinline__ uint32_t segment(const MCDevParams &P, const mcsu::PassConst &kc, float *__restrict__ nx,
                                            float *__restrict__ ny, float *__restrict__ nz, int q, int j, int k,
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
  float4 c[3];
  mcsu::PairNbrs nb;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    c[d] = *reinterpret_cast<const float4 *>(pl[d] + own);
    const float4 yp = *reinterpret_cast<const float4 *>(pl[d] + r_jp);
    const float4 ym = *reinterpret_cast<const float4 *>(pl[d] + r_jm);
    const float4 zp = *reinterpret_cast<const float4 *>(pl[d] + r_kp);
    const float4 zm = *reinterpret_cast<const float4 *>(pl[d] + r_km);
    nb.yp[d] = make_float2(sel(yp, PAR), sel(yp, PAR + 2));
    nb.ym[d] = make_float2(sel(ym, PAR), sel(ym, PAR + 2));
    nb.zp[d] = make_float2(sel(zp, PAR), sel(zp, PAR + 2));
    nb.zm[d] = make_float2(sel(zm, PAR), sel(zm, PAR + 2));
    nb.ov[d] = pl[d][out];
  }
  const int sA = own + PAR;  // linear index of the first site
  mcsu::Site S[2];
  bool a[2];
  MCU4 r;
  mcsu::pair_update<POT, PAR>(P, kc, c, nb, (uint32_t)sA, S, a, r);
  if (a[0] || a[1]) {
    mcsu::pair_apply<PAR>(c, S, a);
#pragma unroll
    for (int d = 0; d < 3; ++d) *reinterpret_cast<float4 *>(pl[d] + own) = c[d];
  }
  if (RECORD) mcsu::pair_record<POT>(P, kc, rec, sA, S, a, r);
  return (a[0] ? 1u : 0u) + (a[1] ? 1u : 0u);
}

template <int POT, bool RECORD>
__global__ __launch_bounds__(kBlock) void k_sweep_x4(MCDevParams P, float *__restrict__ nx, float *__restrict__ ny,
                                                     float *__restrict__ nz, MCDevState *st, int colour, int finalize,
                                                     int xcd, MCRecord rec) {
  const mcsu::PassConst kc = mcsu::pass_const(P, st);
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
      acc += segment<POT, RECORD, 0>(P, kc, nx, ny, nz, q, j, k, rec);
    else
      acc += segment<POT, RECORD, 1>(P, kc, nx, ny, nz, q, j, k, rec);
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
    const mcsu::PassConst kc = mcsu::pass_const(P, st);
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
        acc = segment<POT, false, 0>(P, kc, nx + off, ny + off, nz + off, q, j, k, MCRecord{});
      else
        acc = segment<POT, false, 1>(P, kc, nx + off, ny + off, nz + off, q, j, k, MCRecord{});
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
