// simulator and SimulatorFactory of the host drop-in (reference
// include/simulator.h, include/simulator_factory.h, src/simulatorGPU.cpp).
//
// SimulatorFactory is declared but never defined in the reference; here it is
// implemented and simulator::Setup_simmulation is written on top of it.
#pragma once
#include <memory>
#include <vector>

#include "evolve.h"
#include "geometry.h"
#include "parameters.h"

namespace mc {

class SimulatorFactory {
 public:
  // thermalGPU | stepGPU | quenchGPU | electricGPU (case-insensitive) run on
  // the HIP engine; the CPU names thermal | step | quench | electric run on
  // the same engine with the reference CPU sublattice order and CPU snapshot
  // names.  Unknown names: ConfigError 2 (src/simulatorGPU.cpp:46-49).
  static Evolve *createEvolve(float *ni, int *pt, Parameters *params);
  // bulk | slab | sphere | custom; ConfigError 2 otherwise.
  static Geometry *createGeometry(int *pt, Parameters *params);
  // Re-resolves the axis boundary types after the geometry forced some.
  static void setupBoundaries(Parameters &params, Geometry *geo);
  // Validates the potential name and fills the GHRL coefficients.
  static void setupPotential(Parameters &params, Evolve *ev);
};

class simulator {
 public:
  explicit simulator(Parameters *params);
  ~simulator();
  // src/simulatorGPU.cpp:24-108: evolve mode, geometry, boundaries,
  // anchoring, valid points, initial condition, potential.
  void Setup_simmulation(Parameters &params);
  // ic.csv (src/simulatorGPU.cpp:110-130): "x,y,z,nx,ny,nz,s,pt" with s = 1.
  int print_n(const char *fname, Parameters *params);
  Parameters *params;
  int Nx, Ny, Nz;
  Evolve *evolve = nullptr;
  int *pt = nullptr;
  float *ni = nullptr;

 private:
  std::vector<float> ni_store;
  std::vector<int> pt_store;
  std::unique_ptr<Geometry> geometry_store;
};

}  // namespace mc
