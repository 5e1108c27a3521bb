// Initial director fields of the host drop-in (reference src/ic.cpp).
#pragma once
#include <cstdint>

#include "parameters.h"

namespace mc {

// apply_Initial_Condidions (src/ic.cpp:12-25): random | homogeneous | ic_file
// | cholesteric; ConfigError 2 for an unknown name.  ni is AoS float[3*Nt];
// inactive sites (pt == 0) get (0, 0, 1).
void apply_Initial_Condidions(float *ni, const int *pt, const Parameters &params);

// Directors uniform on the sphere by the Archimedes construction
// (z = 2u - 1, phi = 2*pi*v) with u, v the top 24 bits of Philox4x32-10 words
// 0 and 1 at counter (site, 0, 0, 0x4943) under key (seed, 0).  Replaces GSL
// taus + gsl_ran_dir_3d (src/ic.cpp:27-59): same distribution, own stream.
void random_ic(float *ni, const int *pt, int Nx, int Ny, int Nz, uint64_t seed);

}  // namespace mc
