// Initial director fields.
#include "ic.h"

#include <math.h>
#include <stdio.h>
#include <string.h>
#include <strings.h>

#include <string>

namespace mc {
namespace {

// Philox4x32-10 (Salmon et al., SC'11), host copy for initial conditions.
void philox(uint32_t c[4], uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    c[1] = (uint32_t)p1;
    c[3] = (uint32_t)p0;
    c[0] = n0;
    c[2] = n2;
  }
}

inline void inactive(float *v) {
  v[0] = 0;
  v[1] = 0;
  v[2] = 1;
}

}  // namespace

void random_ic(float *ni, const int *pt, int Nx, int Ny, int Nz, uint64_t seed) {
  const long long nt = (long long)Nx * Ny * Nz;
#pragma omp parallel for schedule(static)
  for (long long s = 0; s < nt; ++s) {
    float *v = ni + 3 * s;
    if (pt && !pt[s]) {
      inactive(v);
      continue;
    }
    uint32_t c[4] = {(uint32_t)s, 0u, 0u, 0x4943u};
    philox(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const double z = 2.0 * ((double)(c[0] >> 8) * (1.0 / 16777216.0)) - 1.0;
    const double phi = 2.0 * M_PI * ((double)(c[1] >> 8) * (1.0 / 16777216.0));
    const double r = sqrt(fmax(0.0, 1.0 - z * z));
    v[0] = (float)(r * cos(phi));
    v[1] = (float)(r * sin(phi));
    v[2] = (float)z;
  }
}

// src/ic.cpp:61-86
static void homogeneous_ic(float *ni, const int *pt, const Parameters &p) {
  const long long nt = (long long)p.Nx * p.Ny * p.Nz;
  const float theta = p.theta_0, phi = p.phi_0;
  printf("  Initial condition: homogeneous\n  %-10s %g\n  %-10s %g\n\n", "phi0:", phi, "theta0:", theta);
  const float v0 = (float)(sin(theta * M_PI / 180) * cos(phi * M_PI / 180));
  const float v1 = (float)(sin(theta * M_PI / 180) * sin(phi * M_PI / 180));
  const float v2 = (float)cos(theta * M_PI / 180);
  for (long long s = 0; s < nt; ++s) {
    if (!pt[s]) {
      inactive(ni + 3 * s);
      continue;
    }
    ni[3 * s] = v0;
    ni[3 * s + 1] = v1;
    ni[3 * s + 2] = v2;
  }
}

// src/ic.cpp:121-151: a helix along z with pitch p0_i (or p0).
static void cholesteric_ic(float *ni, const int *pt, const Parameters &p) {
  const float theta = p.theta_0;
  const float phi_0 = (float)(p.phi_0 * M_PI / 180);
  const float p0 = p.p0_i != 0 ? p.p0_i : p.p0;
  printf("  Initial condition: cholesteric\n  %-10s %g\n  %-10s %g\n  %-10s %g\n\n", "p0:", p0, "phi0:", phi_0,
         "theta0:", theta);
  for (int i = 0; i < p.Nx; ++i)
    for (int j = 0; j < p.Ny; ++j)
      for (int k = 0; k < p.Nz; ++k) {
        const long long s = (long long)i + (long long)p.Nx * (j + (long long)p.Ny * k);
        if (!pt[s]) {
          inactive(ni + 3 * s);
          continue;
        }
        const float phi = (float)(phi_0 - 2 * M_PI * k / p0);
        ni[3 * s] = (float)(sin(theta * M_PI / 180) * cos(phi));
        ni[3 * s + 1] = (float)(sin(theta * M_PI / 180) * sin(phi));
        ni[3 * s + 2] = (float)cos(theta * M_PI / 180);
      }
}

// src/ic.cpp:88-119: the first six columns x,y,z,nx,ny,nz of a CSV with a
// header line (a director_field_*.csv snapshot qualifies).
static void read_ic_file(float *ni, const Parameters &p) {
  FILE *f = fopen(p.ic_file, "r");
  if (!f) throw ConfigError{2, std::string(p.ic_file) + ": cannot open initial-condition file"};
  char buf[512];
  if (!fgets(buf, sizeof(buf), f)) {
    fclose(f);
    throw ConfigError{3, "empty initial-condition file"};
  }
  printf("  Initial condition: file (%s)\n", p.ic_file);
  const long long nt = (long long)p.Nx * p.Ny * p.Nz;
  for (long long n = 0; n < nt; ++n) {
    int x, y, z;
    float a, b, c;
    if (fscanf(f, "%d,%d,%d,%f,%f,%f", &x, &y, &z, &a, &b, &c) != 6 || x < 0 || y < 0 || z < 0 || x >= p.Nx ||
        y >= p.Ny || z >= p.Nz) {
      fclose(f);
      throw ConfigError{3, "Malformed data: problem reading line " + std::to_string(n + 2)};
    }
    const long long s = (long long)x + (long long)p.Nx * (y + (long long)p.Ny * z);
    ni[3 * s] = a;
    ni[3 * s + 1] = b;
    ni[3 * s + 2] = c;
    if (!fgets(buf, sizeof(buf), f) && n + 1 < nt) {
      fclose(f);
      throw ConfigError{3, "Malformed data: file ends at line " + std::to_string(n + 2)};
    }
  }
  fclose(f);
}

void apply_Initial_Condidions(float *ni, const int *pt, const Parameters &p) {
  if (strcasecmp(p.ic, "random") == 0) {
    printf("  Initial condition: random (seed %d)\n\n", p.seed);
    random_ic(ni, pt, p.Nx, p.Ny, p.Nz, (uint64_t)(uint32_t)p.seed);
  } else if (strcasecmp(p.ic, "homogeneous") == 0) {
    homogeneous_ic(ni, pt, p);
  } else if (strcasecmp(p.ic, "ic_file") == 0) {
    read_ic_file(ni, p);
  } else if (strcasecmp(p.ic, "cholesteric") == 0) {
    cholesteric_ic(ni, pt, p);
  } else {
    throw ConfigError{2, std::string("Initial condition '") + p.ic +
                             "' not implemented. Options: random | homogeneous | ic_file | cholesteric"};
  }
}

}  // namespace mc
