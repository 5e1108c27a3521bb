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

// Everything a launch needs about one lattice.
struct MCLaunch {
  MCDevParams P;
  hipStream_t stream;
  float *nx, *ny, *nz;     // SoA directors
  const uint8_t *pt;       // point types, nullptr for an all-bulk lattice
  const float *sx, *sy, *sz;  // SoA surface normals (nullptr when unused)
  MCDevState *st;
};

int mc_set_slots(const MCDevParams &P, int set);
void mc_launch_sweep_pass(const MCLaunch &L, int set, int grid, int finalize, const MCRecord *rec);
void mc_launch_pass_inject(const MCLaunch &L, int set, int grid, const float *prop, const float *u, uint8_t *acc,
                           float *eo, float *en);
// Fast pass for all-active periodic lattices with Nx % 4 == 0 (mc_sweep_x4.hip).
bool mc_x4_eligible(const MCLaunch &L);
int mc_x4_slots(const MCDevParams &P);
void mc_launch_sweep_x4(const MCLaunch &L, int colour, int grid, int finalize, const MCRecord *rec);
void mc_launch_site_energy(const MCLaunch &L, int grid, float *out);
// out = {E, S, q6, n_active}; series (may be null) receives {E, S}.
void mc_launch_measure(const MCLaunch &L, int grid, double *partials, double *out, double *series);
void mc_launch_aos_to_soa(hipStream_t s, const float *aos, float *x, float *y, float *z, int n, int grid);
void mc_launch_soa_to_aos(hipStream_t s, const float *x, const float *y, const float *z, float *aos, int n,
                          int grid);
