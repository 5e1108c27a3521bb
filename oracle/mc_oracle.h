/*
 * CPU oracle for the McLiCS checkerboard Metropolis sweep.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference's
 * CPU path (src/evolve.cpp, src/geometry*.cpp, src/potential.cpp,
 * src/anchoring_*.cpp, src/parameter_order.cpp, src/io.cpp:setGHRL).  It is
 * the checker the HIP library is compared against; only tests/, the smoke
 * check and bench.py's cpu_baseline leg may load it.  Nothing in the product
 * path calls it.
 *
 * Floating-point semantics follow the reference exactly: FP32 dot products
 * evaluated left to right without FMA contraction (built with gcc -O3 and no
 * -march, as Makefile:13 of the reference), the LL/GHRL/pear bond terms with
 * their double literals, the 6-bond sum accumulated in double
 * (src/geometry.cpp:21), and the accept test u < expf(-dE/T)
 * (src/evolve.cpp:160).
 *
 * Arrays use the reference layout: directors AoS float[3*Nt] at
 * (i + Nx*(j + Ny*k))*3 + c (include/define.h:6-12), point types int[Nt]
 * (0 inactive, 1 bulk, >=2 surface index + 2), surface normals AoS float[3*Nt].
 */
#ifndef MC_ORACLE_H_
#define MC_ORACLE_H_
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_FREE = 0, ORC_PERIODIC = 1 };
enum { ORC_LL = 0, ORC_GHRL = 1, ORC_PEAR = 2 };
/* Anchoring kinds, src/geometry.cpp:82-97. */
enum {
  ORC_HOMEOTROPIC = 0,
  ORC_RP = 1,
  ORC_FG = 2,
  ORC_STRONG = 3,
  ORC_HOMEOTROPIC_GHRL = 4,
  ORC_RP_GHRL = 5,
  ORC_FG_GHRL = 6,
  ORC_STRONG_GHRL = 7
};
#define ORC_MAX_SURF 8

typedef struct {
  int32_t nx, ny, nz;
  int32_t bound[3];
  int32_t potential;
  float A, B1, B2, C;
  float ghrl_rho, ghrl_lambda, ghrl_mu, ghrl_nu, ghrl_sigma;
  float elecX, elecY, elecZ, elecA, elecE;
  int32_t n_surf;
  int32_t surf_kind[ORC_MAX_SURF];
  float surf_W[ORC_MAX_SURF];
  float surf_dir[ORC_MAX_SURF][3]; /* unit vector of (theta_s, phi_s) for rp/strong kinds */
} orc_params;

/* Pair, surface and field terms (src/potential.cpp, src/anchoring_*.cpp). */
float orc_bond(const orc_params *p, const float ni[3], const float nj[3], const float r[3]);
float orc_surface(const orc_params *p, int surf, const float ni[3], const float s[3]);
float orc_field(const orc_params *p, const float ni[3]);

/* GHRL coefficients from k11/k22/k33/p0 (src/io.cpp:277-302). */
void orc_set_ghrl(orc_params *p, float k11, float k22, float k33, float p0);

/* Site energy with optional replacement director n0 (NULL = lattice value). */
float orc_site_energy(const orc_params *p, const float *n, const int32_t *pt, const float *ns, int i, int j,
                      int k, const float *n0);
void orc_site_energies(const orc_params *p, const float *n, const int32_t *pt, const float *ns, float *out);

/*
 * One pass over a site set with injected proposals and uniforms.
 *   set = 0..7 : 2x2x2 sublattice tick (di,dj,dk) = (set%2, (set/2)%2, set/4), src/evolve.cpp:100-102
 *   set = 8,9  : checkerboard colour (i+j+k)%2 == set-8
 * prop is AoS float[3*Nt], u float[Nt]; acc/e_old/e_new are written only at
 * the sites of the set.  Returns the number of accepted moves.
 */
int64_t orc_pass_inject(const orc_params *p, float *n, const int32_t *pt, const float *ns, int set, float T,
                        const float *prop, const float *u, uint8_t *acc, float *e_old, float *e_new);

/* ang_var adaptation after a sweep (src/evolve.cpp:169-178). */
float orc_adapt_ang_var(float ang_var, int64_t accepted, int64_t valid);

/*
 * Whole sweeps with the counter-based generator, CPU semantics (no
 * renormalisation, expf accept).  order 0 = 8 sublattice passes in the CPU
 * order, 1 = two checkerboard passes.  Philox counter per site-update:
 * (site, sweep, stream, 0x4D43), key (seed_lo, seed_hi).  Adapts *ang_var.
 * nthreads <= 0 uses the OpenMP default.  Returns the accepted count.
 */
int64_t orc_sweep_rng(const orc_params *p, float *n, const int32_t *pt, const float *ns, float T, float *ang_var,
                      uint32_t seed_lo, uint32_t seed_hi, uint32_t sweep, uint32_t stream, int order, int nthreads);

/* Mean site energy.  gpu_semantics=0: sum over all sites / Nt
 * (src/evolve.cpp:27-62); 1: pt==0 sites contribute 0 and divide by the active
 * count (src/evolve.cu:72-75,435 with the CPU j index). FP64 sum. */
double orc_energy_mean(const orc_params *p, const float *n, const int32_t *pt, const float *ns, int gpu_semantics,
                       int nthreads);

/* Q tensor in FP64 over active sites: q6 = {Qxx, Qxy, Qxz, Qyy, Qyz, Qzz}
 * (src/parameter_order.cpp:30-60, accumulated in double instead of float). */
void orc_qtensor(const float *n, const int32_t *pt, int64_t nt, double q6[6]);
/* Eigenvalues of a symmetric 3x3 matrix, descending (closed form). */
void orc_sym3_eigvals(const double q6[6], double ev[3]);
/* Local order parameter of a site and its active von Neumann neighbours
 * (src/parameter_order.cpp:62-113), FP64. */
double orc_local_S(const orc_params *p, const float *n, const int32_t *pt, int i, int j, int k);

/* Number of rows of the float loop for(x=xi; (int)1e6*sign*(x-xf)>=0; x+=dx),
 * sign = (int)(-dx/|dx|) (src/evolve_thermal.cu:69,79).  Writes up to
 * max_rows values to xs. */
int orc_float_loop(float xi, float xf, float dx, float *xs, int max_rows);

/* One Philox4x32-10 block (for known-answer tests). */
void orc_expf_n(const float *x, float *y, int64_t n);
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

#ifdef __cplusplus
}
#endif
#endif
