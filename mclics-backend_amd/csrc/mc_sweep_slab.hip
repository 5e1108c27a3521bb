// Slab kernel: one launch per full checkerboard sweep (both colours) of an
// all-active periodic lattice with Nx % 4 == 0, streaming the lattice once.
//
// Why.  A colour pass must see every site (its own colour for the old value,
// the other colour for the neighbours), so the two-pass sweep moves the
// lattice through the chip twice and writes it at least once.  Here a sweep
// reads the directors from one set of planes (`in`) and writes the swept
// lattice to a second set (`out`); the caller swaps the two afterwards.
//
// Tiles.  A workgroup owns TY full rows (all of x) of CZ consecutive planes:
// rows y0..y0+TY-1, planes z0..z0+CZ-1.  Its threads own one 16-byte x
// segment [4q, 4q+4) of one of R = TY+4 rows (y0-2 .. y0+TY+1, periodic) and
// walk z together:
//   iteration p (z0 <= p < z0+CZ):
//     red update of plane p+1 on rows y0-1 .. y0+TY   (needs old black sites)
//     black update of plane p on rows y0 .. y0+TY-1   (needs the new red sites)
//     store plane p of the owned rows.
// The red updates of rows y0-1, y0+TY and of planes z0-1, z0+CZ belong to
// neighbouring tiles; this tile recomputes them.  A proposal and its accept
// uniform come from a counter-based generator keyed by (site, sweep), and T
// and ang_var are fixed during a sweep, so every recomputation equals the
// owner's.  Only owned sites are counted and stored.
//
// Data.  Each thread keeps its own column of planes p-1 .. p+3 in registers
// (z neighbours).  In-plane neighbours come from LDS half rows: the black
// sites of a row of plane p (old values) and the red sites (new values),
// each X/2 floats per component, double buffered by plane parity, four half
// slots in all:
//   Bbuf(p & 1): black sites of plane p before the sweep (for red updates)
//   Rbuf(p & 1): red sites of plane p after the red half (for black updates)
// Within iteration p the kernel reads Bbuf((p+1)&1) and Rbuf(p&1) and writes
// Rbuf((p+1)&1) and Bbuf(p&1) (black sites of plane p+2); red updates read
// black sites only and black updates red sites only, so one barrier per
// iteration orders everything.
//
// Parity with the two-pass kernels.  Both colours use the pair update of
// mc_site_update.h with the same counters as the segment kernel, so a sweep
// here leaves the lattice, the acceptance count and ang_var exactly as the
// segment kernel's two passes do.
#include <type_traits>

#include "mc_kernels.h"
#include "mc_site_update.h"

namespace {

constexpr int kMaxThreads = 768;
// The shape with a compile-time instantiation: 64 segments (256 sites) per
// row, 8 rows per tile.
constexpr int kFixedRowSegs = 64;
constexpr int kFixedTileRows = 8;

using mcsu::sel;

struct SlabGeom {
  int X, Y, Z;   // lattice
  int XY;        // sites per plane
  int S;         // segments per row (X / 4)
  int R;         // rows held (TY + 4)
  int X2;        // half-row length (X / 2)
  int TY, CZ;
  int zfree, zsurf;  // slab geometry (MCDevParams)
};

// Offsets of one thread that stay fixed while the tile walks z.
struct LaneOff {
  int cs;        // component stride of a half slot (R * X/2 floats)
  int own;       // this segment's two sites in its half row (r * X/2 + 2q)
  int ov0, ov1;  // the x neighbour outside the segment, PAR 0 (x = 4q-1) / 1 (x = 4q+4)
  int rowbase;   // site index of the segment start within a plane (X*y + 4q)
};

// Plane index z in [-2, Z+2) wrapped onto the periodic lattice.
__device__ __forceinline__ int wrap_z(const SlabGeom &g, int z) { return z < 0 ? z + g.Z : (z >= g.Z ? z - g.Z : z); }

// Plane z exists: always for periodic z, inside [0, Z) for a free z.  GEO
// (the lattice may have free z faces and surfaces) is a template parameter
// throughout, so the periodic lattice's code carries none of it.
template <bool GEO>
__device__ __forceinline__ bool has_plane(const SlabGeom &g, int z) {
  return !GEO || !g.zfree || (z >= 0 && z < g.Z);
}

// A plane beyond a free face reads as zero vectors.
template <bool GEO>
__device__ __forceinline__ void load_seg(const float *__restrict__ const in[3], const SlabGeom &g, const LaneOff &L,
                                         int z, float4 v[3]) {
  if (!has_plane<GEO>(g, z)) {
#pragma unroll
    for (int d = 0; d < 3; ++d) v[d] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    return;
  }
  const int off = g.XY * wrap_z(g, z) + L.rowbase;  // < Nt < 2^31
#pragma unroll
  for (int d = 0; d < 3; ++d) v[d] = *reinterpret_cast<const float4 *>(in[d] + off);
}

// Neighbours and surface of plane z (uniform over the workgroup).
template <bool GEO>
__device__ __forceinline__ mcsu::Where where_z(const SlabGeom &g, int z) {
  mcsu::Where w = mcsu::kBulk;
  if (GEO && g.zfree) {
    if (z + 1 >= g.Z) w.on &= ~(1u << 4);
    if (z <= 0) w.on &= ~(1u << 5);
    if (g.zsurf) w.surf = z == 0 ? 0 : (z == g.Z - 1 ? 1 : -1);
  }
  return w;
}

// Stores the two sites of parity par (elements par, par+2) of the segment
// into a half row.
// (Bitwise select, one v_bitop3_b32 with the multiplexer table 0xCA =
// m ? b : a: a plain ?: on the elements is turned into an indexed load from a
// scratch copy of the segment.)
__device__ __forceinline__ float pick(float a, float b, uint32_t mask) {
  return __uint_as_float(__builtin_amdgcn_bitop3_b32(mask, __float_as_uint(b), __float_as_uint(a), 0xCA));
}

__device__ __forceinline__ void put_half(float *slot, const LaneOff &L, int par, const float4 v[3]) {
  const uint32_t mask = 0u - (uint32_t)(par & 1);
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float2 w = make_float2(pick(v[d].x, v[d].y, mask), pick(v[d].z, v[d].w, mask));
    *reinterpret_cast<float2 *>(slot + d * L.cs + L.own) = w;
  }
}

// Updates the pair of colour sites of segment q in local row r of plane z.
// own: the segment (updated in place); zm/zp: the segments of planes z-1 and
// z+1 in this column; oth: the half-row slot holding the other colour of
// plane z.  Returns the accepted count; records when asked.
template <int POT, bool RECORD, int PAR, bool GEO>
__device__ __forceinline__ uint32_t update_pair(const MCDevParams &P, const mcsu::PassConst &kc, const SlabGeom &g,
                                                const LaneOff &L, float4 own[3], const float4 zm[3],
                                                const float4 zp[3], const float *oth, int z, bool counted,
                                                const MCRecord &rec) {
  mcsu::PairNbrs nb;
  const int mo = PAR == 0 ? L.ov0 : L.ov1;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float *sl = oth + d * L.cs;
    nb.yp[d] = *reinterpret_cast<const float2 *>(sl + L.own + g.X2);
    nb.ym[d] = *reinterpret_cast<const float2 *>(sl + L.own - g.X2);
    nb.zp[d] = make_float2(sel(zp[d], PAR), sel(zp[d], PAR + 2));
    nb.zm[d] = make_float2(sel(zm[d], PAR), sel(zm[d], PAR + 2));
    nb.ov[d] = sl[mo];
  }
  const int sA = g.XY * wrap_z(g, z) + L.rowbase + PAR;
  const mcsu::Where w = where_z<GEO>(g, z);
  mcsu::Site S[2];
  bool a[2];
  MCU4 rr;
  mcsu::pair_update<POT, PAR, GEO>(P, kc, own, nb, (uint32_t)sA, w, S, a, rr);
  mcsu::pair_apply<PAR>(own, S, a);
  if (RECORD && counted) mcsu::pair_record<POT, GEO>(P, kc, rec, sA, w, S, a, rr);
  return counted ? (a[0] ? 1u : 0u) + (a[1] ? 1u : 0u) : 0u;
}

template <int POT, bool RECORD, bool GEO>
__device__ __forceinline__ uint32_t update(const MCDevParams &P, const mcsu::PassConst &kc, const SlabGeom &g,
                                           const LaneOff &L, int colour, float4 own[3], const float4 zm[3],
                                           const float4 zp[3], const float *oth, int y, int z, bool counted,
                                           const MCRecord &rec) {
  if (((y + z + colour) & 1) == 0)
    return update_pair<POT, RECORD, 0, GEO>(P, kc, g, L, own, zm, zp, oth, z, counted, rec);
  return update_pair<POT, RECORD, 1, GEO>(P, kc, g, L, own, zm, zp, oth, z, counted, rec);
}

__device__ __forceinline__ void copy3(float4 dst[3], const float4 src[3]) {
#pragma unroll
  for (int d = 0; d < 3; ++d) dst[d] = src[d];
}

// Sum over a workgroup of up to 1024 threads; thread 0 gets the total.
__device__ __forceinline__ uint32_t slab_block_sum(uint32_t v) {
  __shared__ uint32_t wsum[16];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) wsum[wid] = v;
  __syncthreads();
  uint32_t t = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)((blockDim.x + 63) >> 6); ++w) t += wsum[w];
  return t;
}

// SR, TYC: segments per row and tile rows fixed at compile time (0: read
// from the parameters).  Fixed values turn the LDS strides into instruction
// immediates; with SR = 64 a row is one wave, so every row-dependent value
// (role, parity, plane offsets) is wave-uniform and lives in scalar registers.
template <int SR = 0, int TYC = 0>
__device__ __forceinline__ SlabGeom slab_geom(const MCDevParams &P, int TY, int CZ) {
  SlabGeom g;
  g.X = SR ? 4 * SR : P.nx;
  g.Y = P.ny;
  g.Z = P.nz;
  g.S = SR ? SR : P.nx >> 2;
  g.TY = TYC ? TYC : TY;
  g.CZ = CZ;
  g.R = g.TY + 4;
  g.X2 = 2 * g.S;
  g.XY = g.X * P.ny;
  g.zfree = P.zfree;
  g.zsurf = P.zsurf;
  return g;
}

// One tile (rows y0.., planes z0..) of one lattice: the whole sweep of the
// tile described at the top of this file.  Returns this thread's accepted
// count (owned sites only); every thread of the workgroup must call it.
template <int POT, bool RECORD, bool GEO, int SR = 0>
__device__ __forceinline__ uint32_t slab_tile(const MCDevParams &P, const mcsu::PassConst &kc, const SlabGeom &g,
                                              const float *const in[3], float *const out[3], int y0, int z0,
                                              const MCRecord &rec) {
  extern __shared__ float lds[];
  const int CZ = g.CZ;
  // Rows in thread order: the owned rows of even index, those of odd index,
  // the two red halo rows, then the two rows that are only loaded.  A row's
  // index parity fixes which two sites of each segment take a colour, so on
  // short rows (S < 64, several rows to a wave) a wave then updates one
  // parity with one code path, and whole waves skip the updates their rows
  // do not need.  With S = 64 every wave is one row either way.
  // (SR = 64: the row index is the wave index, made visibly uniform.)
  const int q = threadIdx.x % g.S, TY = g.R - 4, ne = (TY + 1) >> 1;
  const int i = SR == 64 ? __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)) : (int)(threadIdx.x / g.S);
  const int r = i < ne ? 2 + 2 * i
                       : (i < TY ? 3 + 2 * (i - ne)
                                 : (i == TY ? 1 : (i == TY + 1 ? g.R - 2 : (i == TY + 2 ? 0 : (i == TY + 3 ? g.R - 1 : i)))));
  const bool active = r < g.R;  // threads beyond S*R only join barriers
  const int y = ((y0 - 2 + r) % g.Y + g.Y) % g.Y;
  const bool red_row = r >= 1 && r <= g.R - 2;
  const bool owned_row = r >= 2 && r <= g.R - 3;
  LaneOff L;
  L.cs = g.R * g.X2;
  L.own = r * g.X2 + 2 * q;
  L.ov0 = r * g.X2 + (q == 0 ? g.X2 - 1 : 2 * q - 1);
  L.ov1 = r * g.X2 + (q == g.S - 1 ? 0 : 2 * q + 2);
  L.rowbase = g.X * y + 4 * q;

  const size_t H = (size_t)3 * g.R * g.X2;  // floats per half slot
  // Bbuf(par) = lds + par*H, Rbuf(par) = lds + (2+par)*H (computed, not
  // indexed from an array, which would live in scratch memory).
  auto Bbuf = [&](int par) { return lds + (size_t)(par & 1) * H; };
  auto Rbuf = [&](int par) { return lds + (size_t)(2 + (par & 1)) * H; };

  float4 P0[3], P1[3], P2[3], P3[3], P4[3], Q[3];
  uint32_t acc = 0;
  // mcx_expf's table in LDS: the rare exact decisions then read it without
  // waiting for the vector memory operations in flight (the prefetched
  // plane, the stores).  The prologue's first barrier publishes it.
  __shared__ uint64_t s_etab[32];
  if (threadIdx.x < 32) s_etab[threadIdx.x] = mcx_expf_table_entry((int)threadIdx.x);
  mcsu::PassConst kt = kc;
  kt.etab = s_etab;

  // Prologue: planes z0-2 .. z0+2; red updates of z0-1 and z0.
  if (active) {
    load_seg<GEO>(in, g, L, z0 - 2, Q);
    load_seg<GEO>(in, g, L, z0 - 1, P0);
    load_seg<GEO>(in, g, L, z0, P1);
    load_seg<GEO>(in, g, L, z0 + 1, P2);
    load_seg<GEO>(in, g, L, z0 + 2, P3);
    put_half(Bbuf((z0 - 1) & 1), L, ((y + z0 - 1 + 1) & 1), P0);
    put_half(Bbuf(z0 & 1), L, ((y + z0 + 1) & 1), P1);
  }
  __syncthreads();
  if (active && red_row) {
    if (has_plane<GEO>(g, z0 - 1))
      update<POT, false, GEO>(P, kt, g, L, 0, P0, Q, P1, Bbuf((z0 - 1) & 1), y, z0 - 1, false, rec);
    acc += update<POT, RECORD, GEO>(P, kt, g, L, 0, P1, P0, P2, Bbuf(z0 & 1), y, z0, owned_row, rec);
    put_half(Rbuf(z0 & 1), L, ((y + z0) & 1), P1);
  }
  __syncthreads();
  if (active) put_half(Bbuf((z0 + 1) & 1), L, ((y + z0 + 1 + 1) & 1), P2);
  __syncthreads();

  // The walk.  ROLE 2: an owned row (red and black updates, plane stores);
  // ROLE 1: a row that is not owned (the red update on red halo rows, no
  // stores); ROLE 0: the role read from the row.  With one row to a wave
  // (SR = 64) the owned-row waves and the others run separate copies of the
  // loop.  Each copy then issues the same memory operations in every
  // iteration, so the wait for the prefetched plane at the top of an
  // iteration counts out the plane stores issued after that plane's loads
  // instead of waiting for them too (vector loads and stores share one
  // in-order counter).
  auto walk = [&](auto role) {
    constexpr int ROLE = decltype(role)::value;
    const bool red = ROLE == 2 || red_row;
    const bool own = ROLE == 2 || (ROLE == 0 && owned_row);
    for (int it = 0; it < CZ; ++it) {
      const int p = z0 + it;
      if (active) {
        // The ring's last slot is loaded in every iteration (plane p+3, and in
        // the last iteration plane p+2 again, which is not used): a
        // conditional load would make the compiler re-initialise the slot in
        // every iteration.
        load_seg<GEO>(in, g, L, it <= CZ - 2 ? p + 3 : p + 2, P4);
        if (it <= CZ - 2) put_half(Bbuf(p & 1), L, ((y + p + 2 + 1) & 1), P3);
        if (red && has_plane<GEO>(g, p + 1)) {
          acc += update<POT, RECORD, GEO>(P, kt, g, L, 0, P2, P1, P3, Bbuf((p + 1) & 1), y, p + 1,
                                          own && it <= CZ - 2, rec);
          if (it <= CZ - 2) put_half(Rbuf((p + 1) & 1), L, ((y + p + 1) & 1), P2);
        }
        if (own) {
          acc += update<POT, RECORD, GEO>(P, kt, g, L, 1, P1, P0, P2, Rbuf(p & 1), y, p, true, rec);
          const int off = g.XY * p + L.rowbase;
#pragma unroll
          for (int d = 0; d < 3; ++d) {  // streaming stores: nothing reads them before the next launch
            typedef float nf4 __attribute__((ext_vector_type(4)));
            const nf4 v = {P1[d].x, P1[d].y, P1[d].z, P1[d].w};
            __builtin_nontemporal_store(v, reinterpret_cast<nf4 *>(out[d] + off));
          }
        }
      }
      __syncthreads();
      copy3(P0, P1);
      copy3(P1, P2);
      copy3(P2, P3);
      copy3(P3, P4);
    }
  };
  if (SR != 64)
    walk(std::integral_constant<int, 0>{});
  else if (owned_row)
    walk(std::integral_constant<int, 2>{});
  else
    walk(std::integral_constant<int, 1>{});
  return acc;
}

template <int POT, bool RECORD, bool GEO, int SR, int TYC>
__global__ __launch_bounds__(kMaxThreads) void k_sweep_slab(MCDevParams P, const float *__restrict__ inx,
                                                            const float *__restrict__ iny,
                                                            const float *__restrict__ inz, float *__restrict__ outx,
                                                            float *__restrict__ outy, float *__restrict__ outz,
                                                            MCDevState *st, int TY, int CZ, int xcd, MCRecord rec) {
  const SlabGeom g = slab_geom<SR, TYC>(P, TY, CZ);
  const float *const in[3] = {inx, iny, inz};
  float *const out[3] = {outx, outy, outz};
  const mcsu::PassConst kc = mcsu::pass_const(P, st);
  const int nyt = g.Y / TY, nzt = g.Z / CZ, ntiles = nyt * nzt;
  int tile = blockIdx.x;
  if (xcd && (ntiles & 7) == 0) tile = (blockIdx.x & 7) * (ntiles >> 3) + (blockIdx.x >> 3);
  const int y0 = (tile % nyt) * TY, z0 = (tile / nyt) * CZ;
  const uint32_t acc = slab_tile<POT, RECORD, GEO, SR>(P, kc, g, in, out, y0, z0, rec);
  const uint32_t bsum = slab_block_sum(acc);
  mcd_finish_sweep(P, st, bsum, true);
}

// Batched replicas: workgroup b sweeps tile wg[b] = {replica, y0, z0} of its
// replica with that replica's tile shape geo[r] = {ty, cz}; the replicas'
// planes sit at rep_off[r] in one plane set (read) and another (written).
// Each replica evolves exactly as alone.  The last of a replica's workgroups
// to finish adapts its ang_var (mcd_finish_replica).
template <int POT>
__global__ __launch_bounds__(kMaxThreads) void k_batch_slab(const MCDevParams *__restrict__ params, MCDevState *states,
                                                            const int4 *__restrict__ wg, const int2 *__restrict__ geo,
                                                            const long long *__restrict__ rep_off,
                                                            const float *__restrict__ inx, const float *__restrict__ iny,
                                                            const float *__restrict__ inz, float *__restrict__ outx,
                                                            float *__restrict__ outy, float *__restrict__ outz) {
  const int4 d = wg[blockIdx.x];
  const int r = d.x;
  const MCDevParams P = params[r];
  MCDevState *st = states + r;
  const int2 sh = geo[r];
  const SlabGeom g = slab_geom(P, sh.x, sh.y);
  const long long off = rep_off[r];
  const float *const in[3] = {inx + off, iny + off, inz + off};
  float *const out[3] = {outx + off, outy + off, outz + off};
  const mcsu::PassConst kc = mcsu::pass_const(P, st);
  const uint32_t acc = slab_tile<POT, false, false>(P, kc, g, in, out, d.y, d.z, MCRecord{});
  const uint32_t bsum = slab_block_sum(acc);
  if (threadIdx.x == 0) mcd_finish_replica(P, st, bsum, (uint32_t)((P.ny / sh.x) * (P.nz / sh.y)));
}

// Persistent batched sweeps: one launch runs nsweeps sweeps of every replica.
// Workgroups take tiles from one counter per sweep; a grid-wide barrier
// separates the sweeps, so a sweep reads the plane set the previous one
// wrote (sweep w reads a when w is even, b when odd, and writes the other).
// sync[0] counts barrier arrivals over the launch, sync[1] is the fault flag
// (left set for the host), sync[2 + w] is the tile counter of sweep w; the
// host zeroes sync[0] and the counters before the launch.  The grid must fit
// the device at once (the host sizes it from the occupancy); even so every
// wait is bounded: a barrier that has not completed after spin_limit polls
// sets the fault flag and its workgroup returns, so the launch always ends.
//
// Tile fetch: thread 0 takes a ticket into one of two LDS slots, a barrier
// publishes it, and every thread reads the same value, so the whole
// workgroup leaves the loop together.  The slot a thread reads is written
// again only two fetches later, with the tile's own barriers in between.
//
// Memory order across workgroups (and across XCDs, whose L2 caches are not
// coherent with each other): the workgroup barrier orders every wave's
// stores before thread 0, whose device-scope release fence then writes them
// back before it arrives; after the barrier has completed, thread 0's
// device-scope acquire fence invalidates the caches the workgroup reads
// through, before any wave reads the planes, the replicas' states (ang_var,
// sweep index) or the tile counters again.  (One release and one acquire per
// workgroup and sweep: a device-scope fence costs an L2 write-back.)  A
// workgroup also leaves the wait when another has already flagged a fault.
__device__ __forceinline__ bool grid_barrier(unsigned int *sync, unsigned int target, unsigned int spin_limit,
                                             int *ok_slot) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __syncthreads();
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __hip_atomic_fetch_add(sync, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    unsigned int n = 0;
    int ok = 1;
    while (__hip_atomic_load(sync, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      if (++n >= spin_limit || __hip_atomic_load(sync + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) {
        __hip_atomic_fetch_or(sync + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ok = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(32);  // about 2,000 cycles: pollers of one word serialise at its channel
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    *ok_slot = ok;
  }
  __syncthreads();
  const bool ok = *ok_slot != 0;
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  return ok;
}

template <int POT>
__global__ __launch_bounds__(kMaxThreads) void k_batch_slab_persist(
    const MCDevParams *__restrict__ params, MCDevState *states, const int4 *__restrict__ wg,
    const int2 *__restrict__ geo, const long long *__restrict__ rep_off, int ntiles, float *ax, float *ay, float *az,
    float *bx, float *by, float *bz, int nsweeps, unsigned int *sync, unsigned int spin_limit) {
  __shared__ int slot[2];
  __shared__ int ok_slot;
  int k = 0;
  for (int w = 0; w < nsweeps; ++w) {
    const bool odd = (w & 1) != 0;
    float *const ix = odd ? bx : ax, *const iy = odd ? by : ay, *const iz = odd ? bz : az;
    float *const ox = odd ? ax : bx, *const oy = odd ? ay : by, *const oz = odd ? az : bz;
    for (;;) {
      if (threadIdx.x == 0)
        slot[k & 1] = (int)__hip_atomic_fetch_add(sync + 2 + w, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __syncthreads();
      const int t = slot[k & 1];
      ++k;
      if (t >= ntiles) break;
      const int4 d = wg[t];
      const int r = d.x;
      const MCDevParams P = params[r];
      MCDevState *st = states + r;
      const int2 sh = geo[r];
      const SlabGeom g = slab_geom(P, sh.x, sh.y);
      const long long off = rep_off[r];
      const float *const in[3] = {ix + off, iy + off, iz + off};
      float *const out[3] = {ox + off, oy + off, oz + off};
      const mcsu::PassConst kc = mcsu::pass_const(P, st);
      const uint32_t acc = slab_tile<POT, false, false>(P, kc, g, in, out, d.y, d.z, MCRecord{});
      const uint32_t bsum = slab_block_sum(acc);
      if (threadIdx.x == 0) mcd_finish_replica(P, st, bsum, (uint32_t)((P.ny / sh.x) * (P.nz / sh.y)));
    }
    if (w + 1 < nsweeps && !grid_barrier(sync, (unsigned int)(w + 1) * gridDim.x, spin_limit, &ok_slot)) return;
  }
}

__global__ void k_persist_reset(unsigned int *sync, int nsweeps) {
  for (int i = threadIdx.x; i < nsweeps + 2; i += blockDim.x)
    if (i != 1) sync[i] = 0u;
}

// ---------------------------------------------------------------------------
// Measurement on the slab layout: mean site energy (each site's energy with
// the reference arithmetic, src/geometry.cpp:19-72: FP32 bonds summed in
// FP64 and rounded, then surface and field) and the FP64 Q tensor, with the
// lattice read once.  A workgroup owns TY full rows of CZ planes and streams
// z: threads hold their 16-byte segment of rows y0-1 .. y0+TY (periodic) for
// planes p-1, p, p+1 in registers (plane p+2 in flight), and plane p's rows
// go through LDS (double buffered by plane parity, one barrier per plane) for
// the y neighbours and the x neighbours beyond the segment.  Partial sums
// {E, Qxx, Qxy, Qxz, Qyy, Qyz, Qzz, n} per workgroup; the one-block finisher
// reduces them in workgroup order.
__device__ __forceinline__ void slab_block_sum8(const double a[8], double *partials) {
  __shared__ double red[8][16];
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
    for (int w = 0; w < (int)((blockDim.x + 63) >> 6); ++w) v += red[threadIdx.x][w];
    partials[(size_t)blockIdx.x * 8 + threadIdx.x] = v;
  }
}

// The tile of TY rows from y0 and CZ planes from z0: this thread's sums in
// a[8] (zero on entry).  Threads past the tile's TY + 2 rows only join the
// barriers.
template <int POT, bool GEO>
__device__ __forceinline__ void measure_slab_tile(const MCDevParams &P, const float *const in[3], float elecE,
                                                  int TY, int CZ, int y0, int z0, float *lds, double a[8]) {
  // Bonds are symmetric bit for bit: bond(i, j, r) == bond(j, i, -r) with
  // the operations of mcd_bond.  Each product commutes; a dot with -r is the
  // exact negation of the dot with r; a difference b - a is the exact
  // negation of a - b; the two-term sums of GHRL's E1 commute.  So n_i.n_j,
  // |n.r|, the GHRL cross term (n_i x n_j).r and the pear (n_j - n_i).r come
  // out the same from either site.  Each x bond inside a segment and each z
  // bond is therefore computed once and used by both of its sites: 4.25
  // bonds per site instead of 6.
  constexpr bool SYM = true;
  const int X = P.nx, Y = P.ny, Z = P.nz, S = X >> 2, R = TY + 2, XY = X * Y;
  // the TY owned rows first, then the two halo rows (see slab_tile)
  const int q = threadIdx.x % S, i = threadIdx.x / S;
  const int r = i < TY ? i + 1 : (i == TY ? 0 : (i == TY + 1 ? R - 1 : i));
  const bool active = r < R, owned = r >= 1 && r <= TY;
  const int y = ((y0 - 1 + r) % Y + Y) % Y;
  const int rowbase = X * y + 4 * q;
  const size_t H = (size_t)3 * R * X;  // floats per buffer
  const double cA = (double)(-P.A) * 1.5;
  const bool zfree = GEO && P.zfree;
  auto load = [&](int z, float4 v[3]) {
    if (zfree && (z < 0 || z >= Z)) {
#pragma unroll
      for (int d = 0; d < 3; ++d) v[d] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      return;
    }
    const int zz = z < 0 ? z + Z : (z >= Z ? z - Z : z);
    const int off = XY * zz + rowbase;
#pragma unroll
    for (int d = 0; d < 3; ++d) v[d] = *reinterpret_cast<const float4 *>(in[d] + off);
  };
  auto bond = [&](int b, const float o[3], const float m[3]) {
    return mcd_bond<POT>(P, cA, b, o[0], o[1], o[2], m[0], m[1], m[2]);
  };
  float4 Pm[3], Pc[3], Pp[3];
  float bzm[4] = {0.0f, 0.0f, 0.0f, 0.0f};  // SYM: -z bonds of the plane's sites
  if (active) {
    load(z0 - 1, Pm);
    load(z0, Pc);
    load(z0 + 1, Pp);
  }
  if (SYM && owned) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const float o[3] = {sel(Pc[0], e), sel(Pc[1], e), sel(Pc[2], e)};
      const float m[3] = {sel(Pm[0], e), sel(Pm[1], e), sel(Pm[2], e)};
      bzm[e] = bond(5, o, m);
    }
  }
  for (int it = 0; it < CZ; ++it) {
    const int p = z0 + it;
    float *B = lds + (size_t)(p & 1) * H;
    if (active) {
#pragma unroll
      for (int d = 0; d < 3; ++d) *reinterpret_cast<float4 *>(B + d * R * X + r * X + 4 * q) = Pc[d];
    }
    __syncthreads();
    float4 Pn[3];
    if (active && it + 1 < CZ) load(p + 2, Pn);  // in flight while plane p is measured
    if (owned) {
      const bool zp_on = !zfree || p + 1 < Z, zm_on = !zfree || p > 0;
      const int surf = GEO && P.zsurf ? (p == 0 ? 0 : (p == Z - 1 ? 1 : -1)) : -1;
      float4 yp[3], ym[3];
      float ol[3], orr[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const float *row = B + d * R * X + r * X;
        yp[d] = *reinterpret_cast<const float4 *>(row + X + 4 * q);
        ym[d] = *reinterpret_cast<const float4 *>(row - X + 4 * q);
        ol[d] = row[q == 0 ? X - 1 : 4 * q - 1];
        orr[d] = row[q == S - 1 ? 0 : 4 * q + 4];
      }
      // SYM: bx[e] is the bond between x neighbours e-1 and e of the segment
      // (bx[0] with the site left of it, bx[4] with the site right of it);
      // bzp[e] the +z bond of site e
      float bx[5], bzp[4];
      if (SYM) {
#pragma unroll
        for (int e = 0; e <= 4; ++e) {
          float o[3], m[3];
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            o[d] = e == 4 ? sel(Pc[d], 3) : sel(Pc[d], e);
            m[d] = e == 0 ? ol[d] : (e == 4 ? orr[d] : sel(Pc[d], e - 1));
          }
          bx[e] = bond(e == 4 ? 0 : 1, o, m);  // site e's -x bond; site 3's +x bond for e = 4
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const float o[3] = {sel(Pc[0], e), sel(Pc[1], e), sel(Pc[2], e)};
          const float m[3] = {sel(Pp[0], e), sel(Pp[1], e), sel(Pp[2], e)};
          bzp[e] = bond(4, o, m);
        }
      }
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        float o[3], m[6][3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          o[d] = sel(Pc[d], e);
          m[0][d] = e == 3 ? orr[d] : sel(Pc[d], e + 1);  // +x
          m[1][d] = e == 0 ? ol[d] : sel(Pc[d], e - 1);   // -x
          m[2][d] = sel(yp[d], e);
          m[3][d] = sel(ym[d], e);
          m[4][d] = sel(Pp[d], e);
          m[5][d] = sel(Pm[d], e);
        }
        double so = 0.0;
#pragma unroll
        for (int b = 0; b < 6; ++b) {
          if ((b == 4 && !zp_on) || (b == 5 && !zm_on)) continue;
          float v;
          if (SYM && b == 0) v = bx[e + 1];
          else if (SYM && b == 1) v = bx[e];
          else if (SYM && b == 4) v = bzp[e];
          else if (SYM && b == 5) v = bzm[e];
          else v = bond(b, o, m[b]);
          so += (double)v;
        }
        float E = (float)so;
        if (GEO && surf >= 0) E += mcd_surface(P, surf, o[0], o[1], o[2], 0.0f, 0.0f, surf ? 1.0f : -1.0f);
        if (P.elecA != 0.0f) E += mcd_field(P, elecE, o[0], o[1], o[2]);
        // second moments: the product of two FP32 values is exact in FP64
        const double xx = o[0], yy = o[1], zz = o[2];
        a[0] += (double)E;
        a[1] += xx * xx;
        a[2] += xx * yy;
        a[3] += xx * zz;
        a[4] += yy * yy;
        a[5] += yy * zz;
        a[6] += zz * zz;
      }
      a[7] += 4.0;
      if (SYM) {
#pragma unroll
        for (int e = 0; e < 4; ++e) bzm[e] = bzp[e];  // the next plane's -z bonds
      }
    }
    if (!SYM) copy3(Pm, Pc);
    copy3(Pc, Pp);
    if (it + 1 < CZ) copy3(Pp, Pn);
  }
  // the per-site terms of Q: 3 n_a n_b - delta_ab
  a[1] = 3.0 * a[1] - a[7];
  a[4] = 3.0 * a[4] - a[7];
  a[6] = 3.0 * a[6] - a[7];
  a[2] *= 3.0;
  a[3] *= 3.0;
  a[5] *= 3.0;
}

template <int POT, bool GEO>
__global__ __launch_bounds__(kMaxThreads) void k_measure_slab(MCDevParams P, const float *__restrict__ inx,
                                                              const float *__restrict__ iny,
                                                              const float *__restrict__ inz, const MCDevState *st,
                                                              int TY, int CZ, double *__restrict__ partials) {
  extern __shared__ float lds[];
  const int nyt = P.ny / TY, ntiles = nyt * (P.nz / CZ);
  // Workgroup b runs on XCD b % 8: each XCD takes a contiguous run of tiles
  // (neighbouring rows share their halo rows in one L2)
  const int b = (int)blockIdx.x;
  const int tile = (ntiles & 7) == 0 ? (b & 7) * (ntiles >> 3) + (b >> 3) : b;
  const int y0 = (tile % nyt) * TY, z0 = (tile / nyt) * CZ;
  const float *const in[3] = {inx, iny, inz};
  double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  measure_slab_tile<POT, GEO>(P, in, st->elecE, TY, CZ, y0, z0, lds, a);
  slab_block_sum8(a, partials);
}

// Batched replicas (all periodic): workgroup b measures the tile wg[b] =
// {replica, y0, z0} with the replica's sweep tile shape geo[r] = {ty, cz};
// its partial sums go to partials[b], and the replica's workgroups are
// consecutive, so the per-replica finisher sums them in workgroup order.
template <int POT>
__global__ __launch_bounds__(kMaxThreads) void k_batch_measure_slab(const MCDevParams *__restrict__ params,
                                                                    const MCDevState *__restrict__ states,
                                                                    const int4 *__restrict__ wg,
                                                                    const int2 *__restrict__ geo,
                                                                    const long long *__restrict__ rep_off,
                                                                    const float *__restrict__ inx,
                                                                    const float *__restrict__ iny,
                                                                    const float *__restrict__ inz,
                                                                    double *__restrict__ partials) {
  extern __shared__ float lds[];
  const int4 d = wg[blockIdx.x];
  const int r = d.x;
  const MCDevParams P = params[r];
  const int2 sh = geo[r];
  const long long off = rep_off[r];
  const float *const in[3] = {inx + off, iny + off, inz + off};
  double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  measure_slab_tile<POT, false>(P, in, states[r].elecE, sh.x, sh.y, d.y, d.z, lds, a);
  slab_block_sum8(a, partials);
}

template <int POT>
void launch_t(const MCLaunch &L, float *ox, float *oy, float *oz, const MCSlabShape &s, const MCRecord *rec) {
  const int threads = (L.P.nx / 4) * (s.ty + 4);
  const int grid = (L.P.ny / s.ty) * (L.P.nz / s.cz);
  const size_t lds = mc_slab_lds_bytes(L.P, s.ty);
  const MCRecord r = rec ? *rec : MCRecord{};
  // Rows of 256 sites (one wave each) in tiles of 8 rows take the
  // instantiation with that shape fixed at compile time.
  const bool fixed = L.P.nx == 4 * kFixedRowSegs && s.ty == kFixedTileRows;
#define MC_SLAB(REC_, GEO_, SR_, TY_)                                                                          \
  k_sweep_slab<POT, REC_, GEO_, SR_, TY_><<<grid, threads, lds, L.stream>>>(L.P, L.nx, L.ny, L.nz, ox, oy, oz, L.st, \
                                                                            s.ty, s.cz, s.xcd, r)
#define MC_SLAB_S(REC_, GEO_)                            \
  do {                                                   \
    if (fixed)                                           \
      MC_SLAB(REC_, GEO_, kFixedRowSegs, kFixedTileRows); \
    else                                                 \
      MC_SLAB(REC_, GEO_, 0, 0);                         \
  } while (0)
  if (L.P.zfree) {
    if (rec) MC_SLAB_S(true, true);
    else MC_SLAB_S(false, true);
  } else {
    if (rec) MC_SLAB_S(true, false);
    else MC_SLAB_S(false, false);
  }
#undef MC_SLAB_S
#undef MC_SLAB
}

}  // namespace

size_t mc_slab_lds_bytes(const MCDevParams &P, int ty) { return (size_t)4 * 3 * (ty + 4) * (P.nx / 2) * sizeof(float); }

bool mc_slab_shape_ok(const MCDevParams &P, const MCSlabShape &s, size_t lds_limit) {
  if (s.ty < 1 || s.cz < 1 || s.ty > P.ny || s.cz > P.nz) return false;
  if (P.ny % s.ty || P.nz % s.cz) return false;
  if ((P.nx / 4) * (s.ty + 4) > kMaxThreads) return false;
  return mc_slab_lds_bytes(P, s.ty) + 512 <= lds_limit;  // + the static arrays (reduction, exp table)
}

size_t mc_measure_slab_lds_bytes(const MCDevParams &P, int ty) {
  return (size_t)2 * 3 * (ty + 2) * P.nx * sizeof(float);
}

// Whole waves: the workgroup reduction shuffles across all 64 lanes, and the
// lanes beyond the last row hold zero sums.
static int measure_slab_threads(const MCDevParams &P, int ty) { return ((P.nx / 4) * (ty + 2) + 63) / 64 * 64; }

bool mc_measure_slab_shape(const MCDevParams &P, size_t lds_limit, int cu_count, int max_grid, MCSlabShape *out) {
  if ((P.nx % 4) != 0 || !P.per[0] || !P.per[1]) return false;
  // Four-row tiles, at least two per CU: on 256^3 and 128^3 LL this timed
  // about 8 % faster than eight-row tiles with one per CU.
  const int prefs[] = {4, 8, 16, 6, 2, 12, 10, 14, 1};
  for (int ty : prefs) {
    if (ty > P.ny || P.ny % ty) continue;
    if (measure_slab_threads(P, ty) > kMaxThreads) continue;
    if (mc_measure_slab_lds_bytes(P, ty) + 256 > lds_limit) continue;
    int cz = P.nz;
    while (cz % 2 == 0 && cz / 2 >= 4 && (long long)(P.ny / ty) * (P.nz / cz) < 2LL * cu_count) cz /= 2;
    if ((long long)(P.ny / ty) * (P.nz / cz) > max_grid) continue;
    *out = MCSlabShape{ty, cz, 0};
    return true;
  }
  return false;
}

void mc_launch_measure_slab(const MCLaunch &L, const MCSlabShape &s, double *partials, double *out, double *series) {
  const int threads = measure_slab_threads(L.P, s.ty);
  const int grid = (L.P.ny / s.ty) * (L.P.nz / s.cz);
  const size_t lds = mc_measure_slab_lds_bytes(L.P, s.ty);
#define MC_MSLAB(POT_, GEO_) \
  k_measure_slab<POT_, GEO_><<<grid, threads, lds, L.stream>>>(L.P, L.nx, L.ny, L.nz, L.st, s.ty, s.cz, partials)
#define MC_MSLAB_G(POT_)  \
  if (L.P.zfree)          \
    MC_MSLAB(POT_, true); \
  else                    \
    MC_MSLAB(POT_, false)
  switch (L.P.potential) {
    case MCD_GHRL:
      MC_MSLAB_G(MCD_GHRL);
      break;
    case MCD_PEAR:
      MC_MSLAB_G(MCD_PEAR);
      break;
    default:
      MC_MSLAB_G(MCD_LL);
      break;
  }
#undef MC_MSLAB_G
#undef MC_MSLAB
  mc_launch_measure_final(L.stream, partials, grid, out, series);
}

size_t mc_batch_measure_slab_shape(const MCDevParams *hp, const int2 *geo, int nrep, int *threads) {
  size_t lds = 0;
  int t = 0;
  for (int r = 0; r < nrep; ++r) {
    t = std::max(t, measure_slab_threads(hp[r], geo[r].x));
    lds = std::max(lds, mc_measure_slab_lds_bytes(hp[r], geo[r].x));
  }
  *threads = t;
  return lds;
}

void mc_launch_batch_measure_slab(const MCBatchLaunch &B, const MCBatchSlab &S, int threads, size_t lds,
                                  const int *wg0, const int *nwg, double *partials, double *out, double *series) {
#define MC_BMSLAB(POT_)                                                                                           \
  k_batch_measure_slab<POT_><<<S.nwg, threads, lds, B.stream>>>(B.params, B.states, S.wg, S.geo, B.rep_off, B.nx, \
                                                                B.ny, B.nz, partials)
  switch (B.potential) {
    case MCD_GHRL:
      MC_BMSLAB(MCD_GHRL);
      break;
    case MCD_PEAR:
      MC_BMSLAB(MCD_PEAR);
      break;
    default:
      MC_BMSLAB(MCD_LL);
      break;
  }
#undef MC_BMSLAB
  mc_launch_batch_measure_final(B, wg0, nwg, partials, out, series);
}

int mc_slab_prepare(size_t lds_bytes) {
  // Dynamic LDS beyond 64 KiB: ask for it on every instantiation.  The HIP
  // runtime may reject the attribute on devices whose per-workgroup limit is
  // already the full LDS; that is not an error (a launch beyond the limit
  // still fails cleanly and is reported by the caller).
  const void *fns[] = {
      reinterpret_cast<const void *>(&k_batch_slab<MCD_LL>),
      reinterpret_cast<const void *>(&k_batch_slab<MCD_GHRL>),
      reinterpret_cast<const void *>(&k_batch_slab<MCD_PEAR>),
      reinterpret_cast<const void *>(&k_batch_slab_persist<MCD_LL>),
      reinterpret_cast<const void *>(&k_batch_slab_persist<MCD_GHRL>),
      reinterpret_cast<const void *>(&k_batch_slab_persist<MCD_PEAR>),
      reinterpret_cast<const void *>(&k_sweep_slab<MCD_LL, false, false, 0, 0>),
      reinterpret_cast<const void *>(&k_sweep_slab<MCD_LL, false, false, kFixedRowSegs, kFixedTileRows>),
      reinterpret_cast<const void *>(&k_sweep_slab<MCD_LL, true, false, 0, 0>),
      reinterpret_cast<const void *>(&k_sweep_slab<MCD_LL, true, false, kFixedRowSegs, kFixedTileRows>),
      reinterpret_cast<const void *>(&k_sweep_slab<MCD_LL, false, true, 0, 0>),
      reinterpret_cast<const void *>(&k_sweep_slab<MCD_LL, false, true, kFixedRowSegs, kFixedTileRows>),
      reinterpret_cast<const void *>(&k_sweep_slab<MCD_LL, true, true, 0, 0>),
      reinterpret_cast<const void *>(&k_sweep_slab<MCD_LL, true, true, kFixedRowSegs, kFixedTileRows>),
      reinterpret_cast<const void *>(&k_sweep_slab<MCD_GHRL, false, false, 0, 0>),
      reinterpret_cast<const void *>(&k_sweep_slab<MCD_GHRL, false, false, kFixedRowSegs, kFixedTileRows>),
      reinterpret_cast<const void *>(&k_sweep_slab<MCD_GHRL, true, false, 0, 0>),
      reinterpret_cast<const void *>(&k_sweep_slab<MCD_GHRL, true, false, kFixedRowSegs, kFixedTileRows>),
      reinterpret_cast<const void *>(&k_sweep_slab<MCD_GHRL, false, true, 0, 0>),
      reinterpret_cast<const void *>(&k_sweep_slab<MCD_GHRL, false, true, kFixedRowSegs, kFixedTileRows>),
      reinterpret_cast<const void *>(&k_sweep_slab<MCD_GHRL, true, true, 0, 0>),
      reinterpret_cast<const void *>(&k_sweep_slab<MCD_GHRL, true, true, kFixedRowSegs, kFixedTileRows>),
      reinterpret_cast<const void *>(&k_sweep_slab<MCD_PEAR, false, false, 0, 0>),
      reinterpret_cast<const void *>(&k_sweep_slab<MCD_PEAR, false, false, kFixedRowSegs, kFixedTileRows>),
      reinterpret_cast<const void *>(&k_sweep_slab<MCD_PEAR, true, false, 0, 0>),
      reinterpret_cast<const void *>(&k_sweep_slab<MCD_PEAR, true, false, kFixedRowSegs, kFixedTileRows>),
      reinterpret_cast<const void *>(&k_sweep_slab<MCD_PEAR, false, true, 0, 0>),
      reinterpret_cast<const void *>(&k_sweep_slab<MCD_PEAR, false, true, kFixedRowSegs, kFixedTileRows>),
      reinterpret_cast<const void *>(&k_sweep_slab<MCD_PEAR, true, true, 0, 0>),
      reinterpret_cast<const void *>(&k_sweep_slab<MCD_PEAR, true, true, kFixedRowSegs, kFixedTileRows>),
      reinterpret_cast<const void *>(&k_measure_slab<MCD_LL, false>),
      reinterpret_cast<const void *>(&k_measure_slab<MCD_GHRL, false>),
      reinterpret_cast<const void *>(&k_measure_slab<MCD_PEAR, false>),
      reinterpret_cast<const void *>(&k_measure_slab<MCD_LL, true>),
      reinterpret_cast<const void *>(&k_measure_slab<MCD_GHRL, true>),
      reinterpret_cast<const void *>(&k_measure_slab<MCD_PEAR, true>),
      reinterpret_cast<const void *>(&k_batch_measure_slab<MCD_LL>),
      reinterpret_cast<const void *>(&k_batch_measure_slab<MCD_GHRL>),
      reinterpret_cast<const void *>(&k_batch_measure_slab<MCD_PEAR>),
  };
  int failed = 0;
  for (const void *f : fns)
    if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes) != hipSuccess) {
      (void)hipGetLastError();
      ++failed;
    }
  return failed;
}

void mc_launch_sweep_slab(const MCLaunch &L, float *ox, float *oy, float *oz, const MCSlabShape &s,
                          const MCRecord *rec) {
  switch (L.P.potential) {
    case MCD_GHRL:
      launch_t<MCD_GHRL>(L, ox, oy, oz, s, rec);
      break;
    case MCD_PEAR:
      launch_t<MCD_PEAR>(L, ox, oy, oz, s, rec);
      break;
    default:
      launch_t<MCD_LL>(L, ox, oy, oz, s, rec);
      break;
  }
}

void mc_launch_batch_slab(const MCBatchLaunch &B, const MCBatchSlab &S, float *ox, float *oy, float *oz) {
#define MC_BSLAB(POT_)                                                                                       \
  k_batch_slab<POT_><<<S.nwg, S.threads, S.lds, B.stream>>>(B.params, B.states, S.wg, S.geo, B.rep_off, B.nx, \
                                                            B.ny, B.nz, ox, oy, oz)
  switch (B.potential) {
    case MCD_GHRL:
      MC_BSLAB(MCD_GHRL);
      break;
    case MCD_PEAR:
      MC_BSLAB(MCD_PEAR);
      break;
    default:
      MC_BSLAB(MCD_LL);
      break;
  }
#undef MC_BSLAB
}

int mc_batch_persist_grid(int potential, int threads, size_t lds, int cu_count) {
  const void *fn = potential == MCD_GHRL ? reinterpret_cast<const void *>(&k_batch_slab_persist<MCD_GHRL>)
                   : potential == MCD_PEAR ? reinterpret_cast<const void *>(&k_batch_slab_persist<MCD_PEAR>)
                                           : reinterpret_cast<const void *>(&k_batch_slab_persist<MCD_LL>);
  int per = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, threads, lds) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return per * cu_count;
}

void mc_launch_batch_slab_persist(const MCBatchLaunch &B, const MCBatchSlab &S, float *ox, float *oy, float *oz,
                                  int grid, int nsweeps, unsigned int *sync, unsigned int spin_limit) {
  k_persist_reset<<<1, 64, 0, B.stream>>>(sync, nsweeps);
#define MC_PSLAB(POT_)                                                                                         \
  k_batch_slab_persist<POT_><<<grid, S.threads, S.lds, B.stream>>>(B.params, B.states, S.wg, S.geo, B.rep_off, \
                                                                   S.nwg, B.nx, B.ny, B.nz, ox, oy, oz, nsweeps, \
                                                                   sync, spin_limit)
  switch (B.potential) {
    case MCD_GHRL:
      MC_PSLAB(MCD_GHRL);
      break;
    case MCD_PEAR:
      MC_PSLAB(MCD_PEAR);
      break;
    default:
      MC_PSLAB(MCD_LL);
      break;
  }
#undef MC_PSLAB
}
