// mc_sim: command-line entry point of the host drop-in (reference
// src/main.cpp:63-83).  `mc_sim <paramfile>` removes the previous run's
// outputs, parses the parameters, builds the simulator, writes ic.csv and
// runs the evolution mode.  Configuration errors exit with the reference's
// status codes; engine errors exit with 20 as the reference does on CUDA
// errors.  `mc_sim --dry-run <paramfile>` stops after ic.csv, writes the
// initial field as director_field_dryrun.csv and prints the resolved
// configuration as one JSON line, without touching the GPU.
//
// Checkpoints (no reference counterpart):
//   --checkpoint FILE  write the exact chain state to FILE after every po.dat row;
//   --resume FILE      continue the run FILE was written by: the previous
//                      outputs are kept, ic.csv is not rewritten and po.dat
//                      gets the remaining rows, identical to those of an
//                      uninterrupted run;
//   --max-rows N       stop after N rows of this invocation.
#include <dirent.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <stdexcept>
#include <string>

#include "io.h"
#include "simulator.h"

namespace {

bool has_prefix(const char *s, const char *p) { return strncmp(s, p, strlen(p)) == 0; }
bool has_suffix(const char *s, const char *x) {
  const size_t a = strlen(s), b = strlen(x);
  return a >= b && strcmp(s + a - b, x) == 0;
}

// src/main.cpp:43-61
void clean_previous_output() {
  printf("  Removing output files from previous run... ");
  int removed = 0;
  if (DIR *d = opendir(".")) {
    while (dirent *e = readdir(d)) {
      const char *n = e->d_name;
      if (!strcmp(n, "ic.csv") || !strcmp(n, "po.dat") || (has_prefix(n, "director_field_") && has_suffix(n, ".csv"))) {
        remove(n);
        ++removed;
      }
    }
    closedir(d);
  }
  printf("%d file(s) removed.\n\n", removed);
}

void dry_run_report(const mc::simulator &sim, const Parameters &p) {
  const mc::Geometry *g = sim.evolve->geometry;
  printf("DRYRUN {\"nx\":%d,\"ny\":%d,\"nz\":%d,\"geometry\":\"%s\",\"bound\":[%d,%d,%d],\"potential\":\"%s\","
         "\"valid\":%d,\"n_surf\":%d,\"A\":%.9g,\"ghrl\":[%.9g,%.9g,%.9g,%.9g,%.9g],\"T\":%.9g,\"Ti\":%.9g,"
         "\"Tf\":%.9g,\"dT\":%.9g,\"MCS\":%d,\"MCT\":%d,\"seed\":%d,\"evol\":\"%s\",\"ic\":\"%s\",\"elec\":[%.9g,%.9g,"
         "%.9g,%.9g],\"E\":[%.9g,%.9g,%.9g],\"fn\":%d,\"first_file\":%d,\"surfaces\":[",
         p.Nx, p.Ny, p.Nz, g->name(), g->bound[0], g->bound[1], g->bound[2], p.potential, sim.evolve->VallidPoints,
         (int)g->surfaces.size(), p.A, p.ghrl_rho, p.ghrl_lambda, p.ghrl_mu, p.ghrl_nu, p.ghrl_sigma, p.T, p.Ti, p.Tf,
         p.dT, p.MCS, p.MCT, p.seed, p.evol, p.ic, p.elecX, p.elecY, p.elecZ, p.elecA, p.elecEi, p.elecEf, p.elecdE,
         p.fn, p.first_file);
  for (size_t s = 0; s < g->surfaces.size(); ++s)
    printf("%s{\"kind\":%d,\"type\":\"%s\",\"W\":%.9g,\"n_s\":[%.9g,%.9g,%.9g]}", s ? "," : "", g->surfaces[s].kind,
           g->surfaces[s].type.c_str(), g->surfaces[s].W, g->surfaces[s].n_s[0], g->surfaces[s].n_s[1],
           g->surfaces[s].n_s[2]);
  printf("]}\n");
}

}  // namespace

int main(int argc, char **argv) {
  bool dry = false;
  const char *file = nullptr;
  mc::RunControl control;
  for (int a = 1; a < argc; ++a) {
    const bool has_value = a + 1 < argc;
    if (!strcmp(argv[a], "--dry-run")) {
      dry = true;
    } else if (!strcmp(argv[a], "--checkpoint") && has_value) {
      control.checkpoint = argv[++a];
    } else if (!strcmp(argv[a], "--resume") && has_value) {
      control.resume = argv[++a];
    } else if (!strcmp(argv[a], "--max-rows") && has_value) {
      control.max_rows = atoi(argv[++a]);
      if (control.max_rows < 1) {
        fprintf(stderr, "  [ERROR] --max-rows needs a count >= 1\n");
        return 2;
      }
    } else if (argv[a][0] == '-' && argv[a][1] == '-') {
      fprintf(stderr, "  [ERROR] unknown or incomplete option %s\n", argv[a]);
      return 2;
    } else {
      file = argv[a];
    }
  }
  mc::print_header("McLiCS on MI355X - Monte Carlo Liquid Crystal Simulator", 60, "=");
  printf("\n");
  try {
    const bool resuming = !control.resume.empty() && !dry;
    if (!resuming) clean_previous_output();
    Parameters params = mc::read_input_file(file);
    mc::print_parameters(params);
    mc::simulator sim(&params);
    sim.Setup_simmulation(params);
    if (!resuming) sim.print_n("ic.csv", &params);
    sim.evolve->control = control;
    if (dry) {
      // The snapshot writer on the initial field, so it can be checked
      // without a device.
      mc::print_n("director_field_dryrun.csv", sim.ni, params, sim.pt, sim.evolve->geometry->bound);
      dry_run_report(sim, params);
      return 0;
    }
    sim.evolve->run();
  } catch (const mc::ConfigError &e) {
    fprintf(stderr, "  [ERROR] %s\n", e.message.c_str());
    return e.code;
  } catch (const std::exception &e) {
    fprintf(stderr, "  [ERROR] %s\n", e.what());
    return 20;
  }
  mc::print_header("SIMULATION COMPLETE", 60, "=");
  printf("\n");
  return 0;
}
