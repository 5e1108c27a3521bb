// Lattice geometries and surface anchoring of the host drop-in: point types
// and surface normals per site (reference src/geometry_{bulk,slab,sphere,
// custom}.cpp), boundary types per axis and the anchoring table
// (src/geometry.cpp:73-103, src/anchoring_*.cpp).  The site energy itself is
// evaluated on the GPU; here the geometry only describes the lattice.
#pragma once
#include <memory>
#include <string>
#include <vector>

#include "../../include/mclics_hip.h"
#include "parameters.h"

namespace mc {

struct Anchoring {
  std::string type;   // parameter-file name, e.g. "homeotropic_ghrl"
  int kind = 0;       // mc_anchoring
  float W = 0;
  float n_s[3] = {0, 0, 1};  // unit vector of (theta_s, phi_s) for rp/strong kinds
  const char *name() const;  // the reference's class name string
};

class Geometry {
 public:
  explicit Geometry(Parameters *params);
  virtual ~Geometry() = default;
  int Nx, Ny, Nz;
  int nSurfaces = 0;
  std::vector<Anchoring> surfaces;
  std::vector<float> ns;  // AoS float[3*Nt] surface normals
  int bound[3] = {MC_FREE, MC_FREE, MC_FREE};
  Parameters *params;

  // Builds the anchoring table from params->anchoring_type/W/theta_s/phi_s
  // (src/geometry.cpp:73-103 and the anchoring constructors).
  void Boundary_Init(Parameters *params);
  virtual const char *name() const = 0;

 protected:
  void resolve_bounds(Parameters *params, bool x, bool y, bool z);
};

class Bulk_Geometry : public Geometry {
 public:
  Bulk_Geometry(int *pt, Parameters *params);
  const char *name() const override { return "Bulk"; }
};

class Slab_Geometry : public Geometry {
 public:
  Slab_Geometry(int *pt, Parameters *params);
  const char *name() const override { return "Slab"; }
};

class Sphere_Geometry : public Geometry {
 public:
  Sphere_Geometry(int *pt, Parameters *params);
  const char *name() const override { return "Sphere"; }
};

class Custom_Geometry : public Geometry {
 public:
  Custom_Geometry(int *pt, Parameters *params);
  const char *name() const override { return "Custom"; }
};

// Parses "free"/"periodic" (case-insensitive); ConfigError 2 otherwise.
int parse_boundary(const char *axis, const char *value);

}  // namespace mc
