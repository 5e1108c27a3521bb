// Simulation parameters, the host-side mirror of the reference's Parameters
// (include/parameters.h:11-83): same field names, defaults and meaning, so a
// parameter file and the code reading it behave as they do in McLiCS.  The
// boundary function pointers of the reference become enums resolved by the
// geometry, and nothing here depends on GSL.
#pragma once
#include <map>
#include <string>

struct Parameters {
  // lattice
  char geometry[20] = "bulk";
  int Nx = 16;
  int Ny = 16;
  int Nz = 16;
  char XBoundtype[20] = "free";
  char YBoundtype[20] = "free";
  char ZBoundtype[20] = "free";

  // surfaces, indexed by surface id
  char bound_file_name[100] = "";
  std::map<int, float> theta_s;
  std::map<int, float> phi_s;
  std::map<int, float> W;
  std::map<int, std::string> anchoring_type;

  // potential
  char potential[20] = "ll";
  float A = 1.0f;
  float B1 = 0.04f;
  float B2 = 0.4f;  // parsed, unused (as in the reference)
  float C = 0.3f;
  float Ti = 1.2f;
  float Tf = 0.0f;
  float dT = -0.02f;
  float T = 0.0f;
  float ghrl_rho = 0;
  float ghrl_lambda = 0;
  float ghrl_mu = 0;
  float ghrl_nu = 1;
  float ghrl_sigma = 0;

  // Monte Carlo
  char evol[20] = "thermal";
  int MCS = 100;
  int MCT = 100;
  int fn = 1;
  int first_file = 0;
  int seed = 1;

  // initial condition
  char ic[50] = "random";
  char ic_file[100] = "";
  float phi_0 = 0;
  float theta_0 = 0;
  float p0_i = 0;
  float k11 = 1;
  float k22 = 1;
  float k33 = 1;
  float p0 = 0;

  int neighbourKind = 1;  // only nearest neighbours

  // electric field
  float elecX = 0;
  float elecY = 0;
  float elecZ = 0;
  float elecA = 0;
  float elecEi = 0;
  float elecEf = 0;
  float elecdE = 0;
  float elecE = 1;
};

namespace mc {

// A configuration error carrying the exit status the reference uses for it
// (src/io.cpp:91-92,246-247,215-221,313-320; src/simulatorGPU.cpp:47-48).
struct ConfigError {
  int code;
  std::string message;
};

}  // namespace mc
