// Geometries and anchoring of the host drop-in.
#include "geometry.h"

#include <math.h>
#include <stdio.h>
#include <string.h>
#include <strings.h>

#include <algorithm>
#include <fstream>
#include <sstream>

#include "io.h"

namespace mc {
namespace {

struct KindInfo {
  const char *type;
  int kind;
  const char *cls;  // reference class name (Anchoring::getName)
  bool needs_angles;
};

// src/geometry.cpp:82-97 and the name[] members of include/anchoring.h.
const KindInfo kKinds[] = {
    {"rp", MC_RP, "Rapine Papoular Anchoring", true},
    {"fg", MC_FG, "Founier-Galatola like Anchoring", false},
    {"homeotropic", MC_HOMEOTROPIC, "Homeotropic Anchoring", false},
    {"strong", MC_STRONG, "Strong Anchoring", true},
    {"rp_ghrl", MC_RP_GHRL, "Rapine Papoular Anchoring GHRL", true},
    {"fg_ghrl", MC_FG_GHRL, "Founier-Galatola like Anchoring GHRL", false},
    {"homeotropic_ghrl", MC_HOMEOTROPIC_GHRL, "Homeotropic Anchoring GHRL", false},
    {"strong_ghrl", MC_STRONG_GHRL, "Strong Anchoring GHRL", true},
};

[[noreturn]] void missing_parameter(const char *what, int id, const char *cls) {
  throw ConfigError{1, std::string("Parameter ") + what + " not defined for the boundary condition #" +
                           std::to_string(id) + ". The boundary condition " + cls +
                           " needs the aforementioned parameter defined."};
}

}  // namespace

const char *Anchoring::name() const {
  for (const KindInfo &k : kKinds)
    if (k.kind == kind) return k.cls;
  return "";
}

int parse_boundary(const char *axis, const char *value) {
  if (strcasecmp(value, "free") == 0) return MC_FREE;
  if (strcasecmp(value, "periodic") == 0) return MC_PERIODIC;
  throw ConfigError{2, std::string(axis) + " boundary '" + value + "' not implemented. Options: free | periodic"};
}

Geometry::Geometry(Parameters *p) : Nx(p->Nx), Ny(p->Ny), Nz(p->Nz), params(p) {
  if (Nx < 1 || Ny < 1 || Nz < 1) throw ConfigError{2, "lattice dimensions must be positive"};
  ns.assign(3 * (size_t)Nx * Ny * Nz, 0.0f);
}

void Geometry::resolve_bounds(Parameters *p, bool x, bool y, bool z) {
  if (x) bound[0] = parse_boundary("X", p->XBoundtype);
  if (y) bound[1] = parse_boundary("Y", p->YBoundtype);
  if (z) bound[2] = parse_boundary("Z", p->ZBoundtype);
}

void Geometry::Boundary_Init(Parameters *p) {
  surfaces.clear();
  for (int id = 0; id < nSurfaces; ++id) {
    auto it = p->anchoring_type.find(id);
    if (it == p->anchoring_type.end())
      // The reference ends the program with status 0 here (src/geometry.cpp:78-81).
      throw ConfigError{0, "You must define " + std::to_string(nSurfaces) +
                               " boundaries. Please review your input file."};
    const KindInfo *info = nullptr;
    for (const KindInfo &k : kKinds)
      if (strcasecmp(it->second.c_str(), k.type) == 0) info = &k;
    if (!info) throw ConfigError{2, it->second + " boundary condition is not defined"};
    Anchoring a;
    a.type = info->type;
    a.kind = info->kind;
    auto w = p->W.find(id);
    if (w == p->W.end()) missing_parameter("W", id, info->cls);
    a.W = w->second;
    if (info->needs_angles) {
      auto ph = p->phi_s.find(id);
      if (ph == p->phi_s.end()) missing_parameter("phi_s", id, info->cls);
      auto th = p->theta_s.find(id);
      if (th == p->theta_s.end()) missing_parameter("theta_s", id, info->cls);
      // src/anchoring_rp.cpp:42-45: FP32 degrees-to-radians and float trig.
      const float toPi = (float)(M_PI / 180);
      a.n_s[0] = cosf(toPi * ph->second) * sinf(toPi * th->second);
      a.n_s[1] = sinf(toPi * ph->second) * sinf(toPi * th->second);
      a.n_s[2] = cosf(toPi * th->second);
    }
    surfaces.push_back(a);
    char label[24];
    snprintf(label, sizeof(label), "Surface %d:", id);
    print_field(label, a.name());
    print_field("W:", (double)a.W);
  }
}

// src/geometry_bulk.cpp:57-61: every site active.
Bulk_Geometry::Bulk_Geometry(int *pt, Parameters *p) : Geometry(p) {
  print_field("Geometry:", "Bulk");
  nSurfaces = 0;
  std::fill(pt, pt + (size_t)Nx * Ny * Nz, 1);
  resolve_bounds(p, true, true, true);
}

// src/geometry_slab.cpp:47-67: anchoring planes at z=0 (normal -z, pt 2)
// and z=Nz-1 (normal +z, pt 3); z is forced free.
Slab_Geometry::Slab_Geometry(int *pt, Parameters *p) : Geometry(p) {
  print_field("Geometry:", "Slab");
  nSurfaces = 2;
  for (int j = 0; j < Ny; ++j)
    for (int i = 0; i < Nx; ++i) {
      size_t s0 = (size_t)i + (size_t)Nx * j;
      pt[s0] = 2;
      ns[3 * s0 + 0] = 0;
      ns[3 * s0 + 1] = 0;
      ns[3 * s0 + 2] = -1;
      for (int k = 1; k < Nz - 1; ++k) pt[(size_t)i + (size_t)Nx * (j + (size_t)Ny * k)] = 1;
      size_t s1 = (size_t)i + (size_t)Nx * (j + (size_t)Ny * (Nz - 1));
      pt[s1] = 3;
      ns[3 * s1 + 0] = 0;
      ns[3 * s1 + 1] = 0;
      ns[3 * s1 + 2] = 1;
    }
  snprintf(p->ZBoundtype, sizeof(p->ZBoundtype), "free");
  resolve_bounds(p, true, true, true);
}

// src/geometry_sphere.cpp:28-54: a droplet of radius Nz/2 centred at
// (Nx/2, Ny/2, Nz/2) (integer halves), a shell of width 2 anchoring with the
// radial normal, outside inactive; all axes forced free.
Sphere_Geometry::Sphere_Geometry(int *pt, Parameters *p) : Geometry(p) {
  print_field("Geometry:", "Sphere");
  nSurfaces = 1;
  snprintf(p->XBoundtype, sizeof(p->XBoundtype), "free");
  snprintf(p->YBoundtype, sizeof(p->YBoundtype), "free");
  snprintf(p->ZBoundtype, sizeof(p->ZBoundtype), "free");
  const float Hx = (float)(Nx / 2), Hy = (float)(Ny / 2), Hz = (float)(Nz / 2);
  for (int i = 0; i < Nx; ++i)
    for (int j = 0; j < Ny; ++j)
      for (int k = 0; k < Nz; ++k) {
        const float Rx = i - Hx, Ry = j - Hy, Rz = k - Hz;
        const float r = sqrtf(Rx * Rx + Ry * Ry + Rz * Rz);
        const size_t s = (size_t)i + (size_t)Nx * (j + (size_t)Ny * k);
        if (r < Hz - 1) {
          pt[s] = 1;
        } else if (r < Hz + 1) {
          pt[s] = 2;
          ns[3 * s + 0] = Rx / r;
          ns[3 * s + 1] = Ry / r;
          ns[3 * s + 2] = Rz / r;
        } else {
          pt[s] = 0;
        }
      }
  resolve_bounds(p, true, true, true);
}

// src/geometry_custom.cpp:57-128: a CSV with a header naming at least the
// columns nx, ny, nz and pt; the first three columns are x, y, z.  Sites not
// listed stay bulk (pt 1); the number of surfaces is max(pt) - 1.
Custom_Geometry::Custom_Geometry(int *pt, Parameters *p) : Geometry(p) {
  print_field("Geometry:", "Custom");
  std::ifstream f(p->bound_file_name);
  if (!f) throw ConfigError{2, std::string(p->bound_file_name) + ": cannot open boundary file"};
  std::string line;
  std::getline(f, line);
  std::vector<std::string> cols;
  {
    std::stringstream ss(line);
    std::string c;
    while (std::getline(ss, c, ',')) {
      c.erase(std::remove_if(c.begin(), c.end(), [](char ch) { return ch == ' ' || ch == '\r'; }), c.end());
      cols.push_back(c);
    }
  }
  int pos[4] = {-1, -1, -1, -1};
  const char *want[4] = {"nx", "ny", "nz", "pt"};
  for (int c = 0; c < (int)cols.size(); ++c)
    for (int w = 0; w < 4; ++w)
      if (strcasecmp(cols[c].c_str(), want[w]) == 0) pos[w] = c;
  for (int w = 0; w < 4; ++w)
    if (pos[w] < 0) throw ConfigError{10, std::string("boundary file lacks column ") + want[w]};
  std::fill(pt, pt + (size_t)Nx * Ny * Nz, 1);
  int max_kind = 0, lineno = 0;
  while (std::getline(f, line)) {
    if (line.find_first_not_of(" \r\t") == std::string::npos) continue;
    ++lineno;
    std::vector<double> v;
    std::stringstream ss(line);
    std::string c;
    while (std::getline(ss, c, ',')) v.push_back(atof(c.c_str()));
    if ((int)v.size() < (int)cols.size() || v.size() < 3)
      throw ConfigError{10, "malformed boundary file line " + std::to_string(lineno)};
    const int i = (int)v[0], j = (int)v[1], k = (int)v[2];
    if (i >= Nx || j >= Ny || k >= Nz || i < 0 || j < 0 || k < 0)
      throw ConfigError{10, "Point out of simulation box at line " + std::to_string(lineno)};
    const int kind = (int)v[pos[3]];
    if (kind < 0) throw ConfigError{10, "Negative Point Type found at line " + std::to_string(lineno)};
    const size_t s = (size_t)i + (size_t)Nx * (j + (size_t)Ny * k);
    ns[3 * s + 0] = (float)v[pos[0]];
    ns[3 * s + 1] = (float)v[pos[1]];
    ns[3 * s + 2] = (float)v[pos[2]];
    pt[s] = kind;
    max_kind = std::max(max_kind, kind);
  }
  nSurfaces = std::max(0, max_kind - 1);
  printf("  %d boundary surface(s) found.\n\n", nSurfaces);
  resolve_bounds(p, true, true, true);
}

}  // namespace mc
