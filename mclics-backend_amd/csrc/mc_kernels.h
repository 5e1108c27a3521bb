// Launch interface between the C ABI layer and the sweep kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mc_device.h"

// Per-site capture of one recorded production sweep (device arrays of Nt).
struct MCRecord {
  float *u0, *u1, *u2;
  float *px, *py, *pz;
  uint8_t *acc;
  float *eo, *en;
};

// Sets the message mc_last_error() returns on this thread.
void mc_set_error_message(const char *msg);

// MCDevParams from the public parameter block (shared by mc_create and
// mc_batch_create); returns an error message or nullptr.
struct mc_params;
const char *mc_to_dev_params(const mc_params *p, MCDevParams *P);

// Everything a launch needs about one lattice.
struct MCLaunch {
  MCDevParams P;
  hipStream_t stream;
  float *nx, *ny, *nz;     // SoA directors
  const uint8_t *pt;       // point types, nullptr for an all-bulk lattice
  const float *sx, *sy, *sz;  // SoA surface normals (nullptr when unused)
  MCDevState *st;
};

// Launch description of a handle's lattice (for the learned-proposal code).
struct mc_handle;
int mc_handle_launch(mc_handle *h, MCLaunch *out);

int mc_set_slots(const MCDevParams &P, int set);
void mc_launch_sweep_pass(const MCLaunch &L, int set, int grid, int finalize, const MCRecord *rec);
void mc_launch_pass_inject(const MCLaunch &L, int set, int grid, const float *prop, const float *u, uint8_t *acc,
                           float *eo, float *en);
// Fast pass for all-active periodic lattices with Nx % 4 == 0 (mc_sweep_x4.hip).
bool mc_x4_eligible(const MCLaunch &L);
int mc_x4_slots(const MCDevParams &P);
// Measurement of such a lattice in one launch (mc_sweep_x4.hip); out/series as mc_launch_measure.
void mc_launch_measure_x4(const MCLaunch &L, int grid, int xcd, double *partials, double *out, double *series);
void mc_launch_sweep_x4(const MCLaunch &L, int colour, int grid, int finalize, int xcd, const MCRecord *rec);
// Slab kernel (mc_sweep_slab.hip): one launch per sweep, both colours, reading
// the lattice from L.nx.. and writing it to the o* planes.  Same eligibility
// as the segment kernel; a tile is ty full rows of cz planes.
struct MCSlabShape {
  int ty, cz, xcd;
};
size_t mc_slab_lds_bytes(const MCDevParams &P, int ty);
bool mc_slab_shape_ok(const MCDevParams &P, const MCSlabShape &s, size_t lds_limit);
int mc_slab_prepare(size_t lds_bytes);  // number of instantiations that refused the LDS attribute
// Measurement on the slab layout (mc_sweep_slab.hip): same eligibility as the
// fast kernels; the shape has ty rows and cz planes per workgroup and at most
// max_grid workgroups (the partials buffer).  out/series as mc_launch_measure.
size_t mc_measure_slab_lds_bytes(const MCDevParams &P, int ty);
bool mc_measure_slab_shape(const MCDevParams &P, size_t lds_limit, int cu_count, int max_grid, MCSlabShape *out);
void mc_launch_measure_slab(const MCLaunch &L, const MCSlabShape &s, double *partials, double *out, double *series);
void mc_launch_sweep_slab(const MCLaunch &L, float *ox, float *oy, float *oz, const MCSlabShape &s,
                          const MCRecord *rec);
// Batched replicas (mc_sweep_x4.hip / mc_batch.hip).
struct MCBatchLaunch {
  hipStream_t stream;
  int potential;
  int nrep, ntiles;
  const MCDevParams *params;  // per replica
  MCDevState *states;         // per replica
  const int *tile_rep;        // replica of each tile
  const int *rep_tile0;       // first tile of each replica
  const int *rep_ntiles;      // tiles of each replica
  const long long *rep_off;   // first site of each replica in the planes
  float *nx, *ny, *nz;
  uint32_t *blocks_done;
  int xcd;  // tile schedule (mcd_tiles)
};
void mc_launch_batch_x4(const MCBatchLaunch &B, int colour, int grid, int finalize);
// Per-replica ang_var adaptation and sweep advance after a batched sweep.
void mc_launch_batch_finish(const MCBatchLaunch &B);
// Batched slab sweep (mc_sweep_slab.hip): one launch per sweep of every
// replica, reading B.nx.. and writing o*; ang_var is adapted in the launch.
struct MCBatchSlab {
  const int4 *wg;   // per workgroup {replica, y0, z0, 0}
  const int2 *geo;  // per replica {ty, cz}
  int nwg, threads;
  size_t lds;
};
void mc_launch_batch_slab(const MCBatchLaunch &B, const MCBatchSlab &S, float *ox, float *oy, float *oz);
// Persistent form: nsweeps sweeps in one launch of `grid` workgroups (at most
// what the device holds at once: mc_batch_persist_grid), reading B.nx.. in
// even sweeps and o* in odd ones; sync is a device array of nsweeps + 2 words
// (word 1: fault flag, set when a grid barrier timed out after spin_limit
// polls).  The reset of the other words is queued before the launch.
int mc_batch_persist_grid(int potential, int threads, size_t lds, int cu_count);
void mc_launch_batch_slab_persist(const MCBatchLaunch &B, const MCBatchSlab &S, float *ox, float *oy, float *oz,
                                  int grid, int nsweeps, unsigned int *sync, unsigned int spin_limit);
// Per-replica FP64 {E, Qxx..Qzz, n} partial sums per tile, then per-replica
// reduction in tile order: out[r*8..] = {E_mean, S, q6}.
void mc_launch_batch_measure(const MCBatchLaunch &B, int grid, double *tile_partials, double *out, double *series);
// The per-tile partial sums alone (mc_sweep_x4.hip).
void mc_launch_batch_measure_tiles(const MCBatchLaunch &B, int grid, double *tile_partials);
// Per-replica reduction of partial sums: replica r owns partials first[r] ..
// first[r] + count[r] - 1, summed in that order.
void mc_launch_batch_measure_final(const MCBatchLaunch &B, const int *first, const int *count, const double *partials,
                                   double *out, double *series);
// Batched measurement on the slab tiles (mc_sweep_slab.hip): the launch shape
// for the replicas' tile shapes (returns the LDS bytes, sets *threads), then
// the launch: one workgroup per slab tile, partials[nwg * 8], and the
// per-replica finisher over each replica's consecutive workgroups wg0[r] ..
// wg0[r] + nwg[r] - 1.
size_t mc_batch_measure_slab_shape(const MCDevParams *hp, const int2 *geo, int nrep, int *threads);
void mc_launch_batch_measure_slab(const MCBatchLaunch &B, const MCBatchSlab &S, int threads, size_t lds,
                                  const int *wg0, const int *nwg, double *partials, double *out, double *series);
void mc_launch_site_energy(const MCLaunch &L, int grid, float *out);
// Local order parameter per site (NaN where pt == 0).
void mc_launch_local_S(const MCLaunch &L, int grid, float *out);
// out = {E, S, q6, n_active}; series (may be null) receives {E, S}.
void mc_launch_measure(const MCLaunch &L, int grid, double *partials, double *out, double *series);
// One-block reduction of nparts x 8 partials into out (and series).
void mc_launch_measure_final(hipStream_t stream, const double *partials, int nparts, double *out, double *series);
void mc_launch_aos_to_soa(hipStream_t s, const float *aos, float *x, float *y, float *z, int n, int grid);
void mc_launch_soa_to_aos(hipStream_t s, const float *x, const float *y, const float *z, float *aos, int n,
                          int grid);
