// simulator and SimulatorFactory of the host drop-in.
#include "simulator.h"

#include <stdio.h>
#include <string.h>
#include <strings.h>

#include "ic.h"
#include "io.h"

namespace mc {

Evolve *SimulatorFactory::createEvolve(float *ni, int *pt, Parameters *params) {
  struct Mode {
    const char *name;
    int kind;
    bool cpu;
  };
  static const Mode modes[] = {{"thermalGPU", 0, false}, {"stepGPU", 1, false}, {"quenchGPU", 2, false},
                               {"electricGPU", 3, false}, {"thermal", 0, true},  {"step", 1, true},
                               {"quench", 2, true},       {"electric", 3, true}};
  for (const Mode &m : modes) {
    if (strcasecmp(params->evol, m.name) != 0) continue;
    EvolveNGPU *e = nullptr;
    switch (m.kind) {
      case 0: e = new thermalEvolveNGPU(ni, pt, params); break;
      case 1: e = new stepEvolveNGPU(ni, pt, params); break;
      case 2: e = new quenchEvolveNGPU(ni, pt, params); break;
      default: e = new electricEvolveNGPU(ni, pt, params); break;
    }
    e->cpu_semantics = m.cpu;
    return e;
  }
  throw ConfigError{2, std::string("Evol mode '") + params->evol +
                           "' not implemented. Options: thermal | step | quench | electric (add GPU suffix for GPU runs)"};
}

Geometry *SimulatorFactory::createGeometry(int *pt, Parameters *params) {
  if (strcasecmp(params->geometry, "bulk") == 0) return new Bulk_Geometry(pt, params);
  if (strcasecmp(params->geometry, "slab") == 0) return new Slab_Geometry(pt, params);
  if (strcasecmp(params->geometry, "sphere") == 0) return new Sphere_Geometry(pt, params);
  if (strcasecmp(params->geometry, "custom") == 0) return new Custom_Geometry(pt, params);
  throw ConfigError{2, std::string("Geometry '") + params->geometry +
                           "' not implemented. Options: bulk | slab | sphere | custom"};
}

void SimulatorFactory::setupBoundaries(Parameters &params, Geometry *geo) {
  geo->bound[0] = parse_boundary("X", params.XBoundtype);
  geo->bound[1] = parse_boundary("Y", params.YBoundtype);
  geo->bound[2] = parse_boundary("Z", params.ZBoundtype);
}

void SimulatorFactory::setupPotential(Parameters &params, Evolve *) {
  if (strcasecmp(params.potential, "ll") == 0 || strcasecmp(params.potential, "lebwohl-lahser") == 0) {
    printf("  Potential    : Lebwohl-Lasher\n\n");
  } else if (strcasecmp(params.potential, "ghrl") == 0 || strcasecmp(params.potential, "grun-hess") == 0) {
    setGHRL(params);
    printf("  Potential    : Gruhn-Hess (GHRL)\n\n");
  } else if (strcasecmp(params.potential, "pear") == 0) {
    printf("  Potential    : Splay-bend (Pear)\n\n");
  } else {
    throw ConfigError{2, std::string("Potential '") + params.potential + "' not implemented. Options: ll | ghrl | pear"};
  }
}

simulator::simulator(Parameters *p) : params(p), Nx(p->Nx), Ny(p->Ny), Nz(p->Nz) {}

simulator::~simulator() { delete evolve; }

void simulator::Setup_simmulation(Parameters &p) {
  params = &p;
  Nx = p.Nx;
  Ny = p.Ny;
  Nz = p.Nz;
  const size_t nt = (size_t)p.Nx * p.Ny * p.Nz;
  ni_store.assign(3 * nt, 0.0f);
  pt_store.assign(nt, 0);
  ni = ni_store.data();
  pt = pt_store.data();
  evolve = SimulatorFactory::createEvolve(ni, pt, &p);
  geometry_store.reset(SimulatorFactory::createGeometry(pt, &p));
  evolve->geometry = geometry_store.get();
  SimulatorFactory::setupBoundaries(p, evolve->geometry);
  evolve->geometry->Boundary_Init(&p);
  evolve->check_Points(pt, p);
  apply_Initial_Condidions(ni, pt, p);
  SimulatorFactory::setupPotential(p, evolve);
}

int simulator::print_n(const char *fname, Parameters *p) {
  FILE *out = fopen(fname, "w");
  if (!out) {
    perror(fname);
    return 1;
  }
  fputs("x,y,z,nx,ny,nz,s,pt\n", out);
  for (int k = 0; k < p->Nz; ++k)
    for (int j = 0; j < p->Ny; ++j)
      for (int i = 0; i < p->Nx; ++i) {
        const size_t s = (size_t)i + (size_t)p->Nx * (j + (size_t)p->Ny * k);
        fprintf(out, "%d,%d,%d,%g,%g,%g,1,%d\n", i, j, k, (double)ni[3 * s], (double)ni[3 * s + 1],
                (double)ni[3 * s + 2], pt[s]);
      }
  printf("  Snapshot -> %s\n", fname);
  fflush(stdout);
  fclose(out);
  return 0;
}

}  // namespace mc
