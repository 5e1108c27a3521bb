// Evolution mode drivers of the host drop-in (reference include/evolve.h,
// include/evolve.cuh, src/evolve_{thermal,step,quench,electric}.cu).
//
// Class names and constructor signatures follow the reference so code that
// builds an Evolve from (ni, pt, params) keeps working.  Each run() keeps the
// reference's loop structure, po.dat rows and snapshot names; the lattice
// stays on the GPU for the whole run and only E and S per measurement sweep
// (FP64, reduced on the device) and one snapshot per outer step come back.
#pragma once
#include <stdint.h>

#include <string>

#include "../../include/mclics_hip.h"
#include "geometry.h"
#include "parameters.h"

namespace mc {

// Checkpointing of a run (no reference counterpart: the reference restarts
// from a director snapshot only, src/ic.cpp:88-119).  With `checkpoint` set,
// run() writes the exact state of the chain after every outer row (po.dat
// row); with `resume` set it continues the run such a file was written by;
// max_rows > 0 ends this invocation after that many rows.
struct RunControl {
  std::string checkpoint, resume;
  int max_rows = 0;
};

class Evolve {
 public:
  virtual ~Evolve() = default;
  virtual int run() { return 0; }
  Geometry *geometry = nullptr;
  int VallidPoints = 0;
  RunControl control;
  // src/evolve.cpp:180-193
  void check_Points(int *pt, const Parameters &params);
};

// Base of the GPU modes: owns the engine handle.
class EvolveNGPU : public Evolve {
 public:
  EvolveNGPU(float *ni, int *pt, Parameters *params);
  ~EvolveNGPU() override;

  // The CPU-named modes (thermal, step, quench, electric) run on the same
  // engine with the reference CPU sublattice order and the CPU snapshot names.
  bool cpu_semantics = false;

 protected:
  void open();
  void sweeps(int n);
  // MCS measured sweeps; returns the means of E, E^2, S, S^2 (FP64).
  void measured(int n, double &E, double &E2, double &S, double &S2);
  void set_T(float T);
  void set_field(float E);
  void fetch();
  void snapshot(const std::string &fname);
  FILE *open_po(const char *first_col);
  void row(FILE *f, const char *x, double E, double E2, double S1, double S2);
  double polarization() const;

  // Checkpoints.  start_run() checks control.resume against this run before
  // any device work, opens the handle, restores the chain from the file and
  // sets rows_done and *x (the loop variable of the next row); without a
  // resume file it only opens the handle.  after_row() counts the row,
  // writes control.checkpoint and says whether this invocation stops here.
  void start_run(float *x);
  bool after_row(float next_x);
  void write_checkpoint(float next_x);
  uint64_t chain_fingerprint() const;
  int rows_done = 0;
  int rows_this_run = 0;

  float *ni;
  int *pt;
  Parameters *params;
  int Nx, Ny, Nz;
  mc_handle *h = nullptr;
};

class thermalEvolveNGPU : public EvolveNGPU {
 public:
  using EvolveNGPU::EvolveNGPU;
  int run() override;
};
class stepEvolveNGPU : public EvolveNGPU {
 public:
  using EvolveNGPU::EvolveNGPU;
  int run() override;
};
class quenchEvolveNGPU : public EvolveNGPU {
 public:
  using EvolveNGPU::EvolveNGPU;
  int run() override;
};
class electricEvolveNGPU : public EvolveNGPU {
 public:
  using EvolveNGPU::EvolveNGPU;
  int run() override;
};

// Writes a director snapshot (src/io.cpp:50-80): "x,y,z,nx,ny,nz,S,pt"
// with the local order parameter of each active site and NaN elsewhere.
// S_site, when given, holds those values (from mc_local_order); otherwise
// they are computed here.
int print_n(const char *fname, const float *ni, const Parameters &params, const int *pt, const int bound[3],
            const float *S_site = nullptr);

// Local order parameter of site (i,j,k) and its active von Neumann
// neighbours (src/parameter_order.cpp:62-113), FP64.
double lattice_order_parameter(const float *ni, const int *pt, int i, int j, int k, const Parameters &params,
                               const int bound[3]);

}  // namespace mc
