// Parameter file parser and console helpers of the host drop-in.
//
// The parser reproduces the reference's stream semantics (src/io.cpp:81-253)
// because drop-in users rely on them: values are read with operator>> into
// the field's own type, the rest of the line is discarded with getline, and
// the stream state is checked after every key exactly as check_error_bits
// does (src/io.cpp:306-321): end of file ends parsing quietly, any other
// failure is an invalid value.
#include "io.h"

#include <errno.h>
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <strings.h>

#include <fstream>
#include <functional>
#include <vector>

namespace mc {
namespace {

struct Key {
  std::vector<const char *> names;
  std::function<void(std::ifstream &, Parameters &)> read;
};

template <typename T>
std::function<void(std::ifstream &, Parameters &)> scalar(T Parameters::*field) {
  return [field](std::ifstream &f, Parameters &p) { f >> (p.*field); };
}

template <size_t N>
std::function<void(std::ifstream &, Parameters &)> text(char (Parameters::*field)[N]) {
  return [field](std::ifstream &f, Parameters &p) {
    std::string tok;
    if (f >> tok) {
      strncpy(p.*field, tok.c_str(), N - 1);
      (p.*field)[N - 1] = '\0';
    }
  };
}

template <typename T>
std::function<void(std::ifstream &, Parameters &)> indexed(std::map<int, T> Parameters::*field) {
  return [field](std::ifstream &f, Parameters &p) {
    int id = 0;
    f >> id;
    f >> (p.*field)[id];
  };
}

const std::vector<Key> &keys() {
  static const std::vector<Key> table = {
      {{"nx"}, scalar(&Parameters::Nx)},
      {{"ny"}, scalar(&Parameters::Ny)},
      {{"nz"}, scalar(&Parameters::Nz)},
      {{"A"}, scalar(&Parameters::A)},
      {{"B1"}, scalar(&Parameters::B1)},
      {{"B2"}, scalar(&Parameters::B2)},
      {{"C"}, scalar(&Parameters::C)},
      {{"Ti"}, scalar(&Parameters::Ti)},
      {{"Tf"}, scalar(&Parameters::Tf)},
      {{"dT"}, scalar(&Parameters::dT)},
      {{"MCS"}, scalar(&Parameters::MCS)},
      {{"MCT"}, scalar(&Parameters::MCT)},
      {{"seed"}, scalar(&Parameters::seed)},
      {{"ic"}, text(&Parameters::ic)},
      {{"ic_file"}, text(&Parameters::ic_file)},
      {{"xbound"}, text(&Parameters::XBoundtype)},
      {{"ybound"}, text(&Parameters::YBoundtype)},
      {{"zbound"}, text(&Parameters::ZBoundtype)},
      {{"potential"}, text(&Parameters::potential)},
      {{"evol", "mode"}, text(&Parameters::evol)},
      {{"k11"}, scalar(&Parameters::k11)},
      {{"k22"}, scalar(&Parameters::k22)},
      {{"k33"}, scalar(&Parameters::k33)},
      {{"p0"}, scalar(&Parameters::p0)},
      {{"phi_0"}, scalar(&Parameters::phi_0)},
      {{"theta_0"}, scalar(&Parameters::theta_0)},
      {{"p0_i"}, scalar(&Parameters::p0_i)},
      {{"fn"}, scalar(&Parameters::fn)},
      {{"first_file_number", "initial_output"}, scalar(&Parameters::first_file)},
      {{"geometry"}, text(&Parameters::geometry)},
      {{"boundary_file"}, text(&Parameters::bound_file_name)},
      {{"anchoring_type"}, indexed(&Parameters::anchoring_type)},
      {{"W"}, indexed(&Parameters::W)},
      {{"phi_s"}, indexed(&Parameters::phi_s)},
      {{"theta_s"}, indexed(&Parameters::theta_s)},
      {{"elecX", "Electric_field_x", "EF_x"}, scalar(&Parameters::elecX)},
      {{"elecY", "Electric_field_y", "EF_y"}, scalar(&Parameters::elecY)},
      {{"elecZ", "Electric_field_z", "EF_z"}, scalar(&Parameters::elecZ)},
      {{"Ei", "initial_E"}, scalar(&Parameters::elecEi)},
      {{"Ef", "final_E"}, scalar(&Parameters::elecEf)},
      {{"dE", "E_variation"}, scalar(&Parameters::elecdE)},
      {{"elecA", "dielectric_anisotropy", "Aniso_E"}, scalar(&Parameters::elecA)},
  };
  return table;
}

bool is_nk(const char *tok) {
  return strcasecmp(tok, "nk") == 0 || strcasecmp(tok, "neighbour_kind") == 0 || strcasecmp(tok, "neighbour") == 0;
}

}  // namespace

Parameters read_input_file(const char *fname) {
  Parameters params;
  if (fname == nullptr) {
    printf("  [!] No parameter file specified - using default parameter values.\n\n");
    return params;
  }
  std::ifstream f(fname);
  if (!f) throw ConfigError{5, std::string(fname) + ": " + strerror(errno)};
  printf("  Parameter file : %s\n\n", fname);
  std::vector<char> rest(400);
  std::string tok;
  while (f >> tok) {
    const char *t = tok.c_str();
    bool known = false;
    if (is_nk(t)) {
      int nk = 0;
      f >> nk;
      f.getline(rest.data(), 400);
      if (nk != 1)
        throw ConfigError{1, "neighbour_kind = " + std::to_string(nk) +
                                 " is not supported; only nk = 1 (nearest neighbours) is allowed"};
      params.neighbourKind = 1;
      known = true;
    } else if (t[0] == '#') {
      f.getline(rest.data(), 400);
      known = true;
    } else {
      for (const Key &k : keys()) {
        for (const char *name : k.names) {
          if (strcasecmp(t, name) == 0) {
            k.read(f, params);
            f.getline(rest.data(), 400);
            known = true;
            break;
          }
        }
        if (known) break;
      }
    }
    if (!known) throw ConfigError{3, "Unknown parameter: '" + tok + "'"};
    if (f.eof()) break;
    if (f.fail() || f.bad()) throw ConfigError{1, "Invalid value for parameter: " + tok};
  }
  return params;
}

void setGHRL(Parameters &params) {
  const float k11 = params.k11, k22 = params.k22, k33 = params.k33, p0 = params.p0;
  const float Scale = (float)fabs(9.0 / (double)(k11 - 3 * k22 - k33));
  params.ghrl_sigma = p0 != 0 ? (float)(-Scale * k22 * (2 * M_PI / p0)) : 0.0f;
  params.ghrl_lambda = (Scale / 9) * (2 * k11 - 3 * k22 + k33);
  params.ghrl_mu = Scale * (k22 - k11);
  params.ghrl_rho = (Scale / 9) * (k11 - k33);
  params.ghrl_nu = (Scale / 9) * (k11 - 3 * k22 - k33);
  print_header("ELASTIC PARAMETERS");
  printf("  %-12s %g\n  %-12s %g\n  %-12s %g\n", "k11:", k11, "k22:", k22, "k33:", k33);
  printf("  %-12s %g\n  %-12s %g\n  %-12s %g\n", "lambda:", params.ghrl_lambda, "mu:", params.ghrl_mu, "rho:",
         params.ghrl_rho);
  printf("  %-12s %g\n  %-12s %g\n  %-12s %g\n", "nu:", params.ghrl_nu, "sigma:", params.ghrl_sigma, "scale:", Scale);
  print_separator();
  printf("\n");
}

void print_separator(const char *seg, int count) {
  for (int i = 0; i < count; i++) fputs(seg, stdout);
  putchar('\n');
}

void print_header(const char *title, int count, const char *seg) {
  print_separator(seg, count);
  int pad = (count - (int)strlen(title) - 2) / 2;
  if (pad < 0) pad = 0;
  printf("%*s %s %*s\n", pad, "", title, pad, "");
  print_separator(seg, count);
}

void print_field(const char *label, const char *value) { printf("  %-13s%s\n", label, value); }
void print_field(const char *label, double value) { printf("  %-13s%g\n", label, value); }
void print_field(const char *label, int value) { printf("  %-13s%d\n", label, value); }

void print_parameters(const Parameters &p) {
  print_header("SIMULATION PARAMETERS");
  printf("  %-20s %dx%dx%d\n", "Grid:", p.Nx, p.Ny, p.Nz);
  printf("  %-20s %s\n", "Geometry:", p.geometry);
  printf("  %-20s %s\n", "Potential:", p.potential);
  printf("  %-20s %s\n", "Mode:", p.evol);
  printf("  %-20s %d steps\n", "MCS:", p.MCS);
  printf("  %-20s %d steps\n", "MCT:", p.MCT);
  printf("  %-20s %d\n", "Seed:", p.seed);
  if (p.elecA != 0) printf("  %-20s (%g, %g, %g) x %g\n", "Electric field:", p.elecX, p.elecY, p.elecZ, p.elecA);
  print_separator();
  printf("\n");
}

int format_g(char *buf, double v) { return sprintf(buf, "%g", v); }

}  // namespace mc
