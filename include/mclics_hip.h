/*
 * mclics_hip.h - C ABI of the MI355X (gfx950) checkerboard Metropolis engine.
 *
 * Plain C linkage, plain pointers and sizes, no HIP or C++ types.  Every
 * function returns 0 on success and a negative MC_E* code on failure;
 * mc_last_error() returns the message of the last failure on the calling
 * thread.  Nothing here calls exit().
 *
 * Host arrays handed in use the reference's layout (include/define.h:6-12 of
 * the reference): directors AoS float[3*Nt] at (i + Nx*(j + Ny*k))*3 + c,
 * point types int32[Nt] (0 inactive, 1 bulk, >= 2 surface index + 2) and
 * surface normals AoS float[3*Nt].  On the device the lattice is kept as three
 * FP32 SoA planes.
 */
#ifndef MCLICS_HIP_H_
#define MCLICS_HIP_H_
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MC_MAX_SURF 8

enum mc_status {
  MC_OK = 0,
  MC_EINVAL = -1,  /* bad argument or unsupported configuration */
  MC_EHIP = -2,    /* HIP runtime error (message names the call) */
  MC_ENOMEM = -3,  /* device allocation failed */
  MC_ENODEV = -4,  /* no usable gfx950 device */
  MC_ESTATE = -5   /* call not valid in the handle's current state */
};

enum mc_boundary { MC_FREE = 0, MC_PERIODIC = 1 };
enum mc_potential { MC_LL = 0, MC_GHRL = 1, MC_PEAR = 2 };
enum mc_anchoring {
  MC_HOMEOTROPIC = 0,
  MC_RP = 1,
  MC_FG = 2,
  MC_STRONG = 3,
  MC_HOMEOTROPIC_GHRL = 4,
  MC_RP_GHRL = 5,
  MC_FG_GHRL = 6,
  MC_STRONG_GHRL = 7
};
/* Sweep orders.  MC_ORDER_REF8 visits the 8 sublattices of 2x2x2 cells in the
 * reference CPU order tick -> (tick%2, (tick/2)%2, tick/4)
 * (src/evolve.cpp:100-102); MC_ORDER_CHECKER updates the two (i+j+k) parity
 * colours, a different but equally valid Markov chain. */
enum mc_order { MC_ORDER_CHECKER = 0, MC_ORDER_REF8 = 1 };
/* Sweep kernels.  The segment kernel (X4, one launch per colour) and the slab
 * kernel (SLAB, both colours in one launch that streams the lattice once) need
 * an all-active lattice, periodic on every axis, with Nx % 4 == 0 and the
 * checkerboard order; AUTO picks SLAB when a tile shape fits the device, else
 * X4, else the generic site kernel.  X4 and SLAB draw the same numbers and
 * evolve the lattice identically; the generic kernel uses another counter
 * layout of the generator, so it is a different (equally valid) chain from
 * the same seed. */
enum mc_kernel {
  MC_KERNEL_AUTO = 0,
  MC_KERNEL_GENERIC = 1,
  MC_KERNEL_X4 = 2,
  MC_KERNEL_SLAB = 3,
  MC_KERNEL_SLAB_PERSISTENT = 4 /* batches only: up to 64 sweeps per launch */
};

/* POD flattening of the reference's Parameters (include/parameters.h:11-83)
 * restricted to what the sweep reads.  The std::map surface tables become
 * fixed arrays indexed by surface id. */
typedef struct mc_params {
  int32_t nx, ny, nz;
  int32_t bound[3];          /* mc_boundary per axis */
  int32_t potential;         /* mc_potential */
  float A, B1, B2, C;        /* B2 is parsed by the reference but unused */
  float ghrl_rho, ghrl_lambda, ghrl_mu, ghrl_nu, ghrl_sigma;
  float elecX, elecY, elecZ, elecA, elecE;
  float T;
  int32_t n_surf;
  int32_t surf_kind[MC_MAX_SURF]; /* mc_anchoring */
  float surf_W[MC_MAX_SURF];
  /* Anchoring direction of the rp, strong, rp_ghrl and strong_ghrl kinds: the
   * UNIT VECTOR (sin t cos p, sin t sin p, cos t) of the surface's
   * (theta_s, phi_s) = (t, p) in degrees (src/anchoring_rp.cpp:42-45), not
   * the angles; mc_create rejects a direction of these kinds whose length is
   * not 1 (within 1e-3). */
  float surf_dir[MC_MAX_SURF][3];
  uint64_t seed;                  /* keys the counter-based generator */
  uint32_t stream;                /* independent stream id under one seed */
  int32_t order;                  /* mc_order */
  float ang_var;                  /* initial proposal width, 0.5 in every run() */
} mc_params;

typedef struct mc_handle mc_handle;

const char *mc_last_error(void);
/* Number of visible gfx950 devices (0 when none); never fails. */
int mc_device_count(void);

/* Replaces the device allocation and upload done at the start of every
 * reference run(): src/evolve_thermal.cu:45-64 (and src/evolve_step.cu:42-64,
 * src/evolve_quench.cu:42-64, src/evolve_electric.cu:45-65).  ns may be NULL when no site has
 * pt >= 2. */
int mc_create(const mc_params *params, const float *ni_aos, const int32_t *pt, const float *ns_aos,
              mc_handle **out);
int mc_destroy(mc_handle *h);

/* Temperature and field strength as read by the next sweep
 * (src/evolve_thermal.cu:80 H2D of T; src/evolve_electric.cu:82 resend of
 * Parameters with a new elecE). */
int mc_set_T(mc_handle *h, float T);
int mc_set_field(mc_handle *h, float elecE);
int mc_set_ang_var(mc_handle *h, float ang_var);
int mc_get_ang_var(mc_handle *h, float *ang_var);

/* Sweep kernel selection (mc_kernel); get returns the kernel in use. */
int mc_set_sweep_kernel(mc_handle *h, int kind);
int mc_get_sweep_kernel(mc_handle *h, int *kind);

/* Launch shape of the segment kernel: grid = number of persistent
 * workgroups (0 = 8 per CU), xcd = 1 gives each of the 8 XCDs a contiguous
 * eighth of the tiles (grid must then be a multiple of 8). */
int mc_set_tuning(mc_handle *h, int grid, int xcd);
/* Tile shape of the slab kernel: ty full rows (Ny % ty == 0) of cz planes
 * (Nz % cz == 0) per workgroup, xcd = 1 keeps the tiles of one group of
 * planes on one XCD.  0 for ty or cz restores the default shape.  Rejected
 * (MC_EINVAL) when the shape does not fit the device. */
int mc_set_slab_tile(mc_handle *h, int ty, int cz, int xcd);
int mc_get_slab_tile(mc_handle *h, int *ty, int *cz, int *xcd);
/* Times the candidate launch shapes of the kernel in use (segment: grid and
 * xcd; slab: tile shape, reported as grid = number of workgroups) on a
 * scratch copy of the lattice (the chain itself is not advanced), keeps the
 * fastest and reports it with its mean sweep time in microseconds (any output
 * may be NULL). */
int mc_autotune(mc_handle *h, int *grid, int *xcd, float *best_us);

/* n full sweeps (EvolveNGPU::Monte_Carlo_Step_GPU, src/evolve.cu:265-388),
 * asynchronous on the handle's stream, with the ang_var adaptation done on
 * the device after each sweep so no host round trip is needed. */
int mc_sweeps(mc_handle *h, int n);
/* Same, through captured hipGraphs of at most 64 sweeps each: the graph of
 * 64 sweeps is replayed n / 64 times, then the graph of the n % 64 that
 * remain.  Graphs are captured on first use and cached per handle (at most
 * 128 of them); the chain equals mc_sweeps's bit for bit.  With a learned
 * proposal attached, sweeps are captured when every sweep is learned
 * (every = 1) and run eagerly otherwise; attaching a network or switching
 * its reflection move drops the cached graphs. */
int mc_sweeps_graph(mc_handle *h, int n);
/* Blocks until queued work is done. */
int mc_sync(mc_handle *h);
/* Runs n sweeps bracketed by HIP events on the handle's stream; returns the
 * elapsed milliseconds and the number of sweep-kernel launches made. */
int mc_time_sweeps(mc_handle *h, int n, int use_graph, float *ms, int *launches);

/* Runs n sweeps eagerly between one HIP event pair on the handle's stream
 * and returns the region's elapsed time and the number of pass launches.
 * With every >= 1, every `every`-th pass launch (the first included) also
 * gets an event pair of its own and avg_pass_ms is the mean of those.  With
 * every = 0 no pass is bracketed, the launches run back to back as in a run,
 * and avg_pass_ms is the region's time over its pass launches (the kernels
 * plus the gaps between them; an event pair around a single launch adds a
 * few microseconds of its own). */
int mc_time_passes(mc_handle *h, int n, int every, double *avg_pass_ms, double *total_ms, int *passes);

/* Mean site energy and order parameter, FP64 reductions on the device
 * (energy_calculator_GPU src/evolve.cu:390-436 plus Matrice_constructor /
 * Eigen_value_evaluation src/parameter_order.cpp:11-60).  E is the sum of
 * FP32 site energies over active sites divided by their count; q6 (may be
 * NULL) receives {Qxx,Qxy,Qxz,Qyy,Qyz,Qzz}. */
int mc_measure(mc_handle *h, double *E, double *S, double *q6);
/* n times (one sweep, then the measurement above), queued without host
 * synchronisation; E_series/S_series (double[n], either may be NULL) receive
 * the per-sweep values.  This is the MCS loop of the mode drivers
 * (src/evolve_thermal.cu:85-95) with the lattice never leaving the device. */
int mc_measure_series(mc_handle *h, int n, double *E_series, double *S_series);
/* The same with `every` sweeps before each measurement (every = 1 is
 * mc_measure_series): thinned series for long autocorrelation studies.  No
 * reference counterpart: the reference measures after every MCS sweep. */
int mc_measure_series_every(mc_handle *h, int n, int every, double *E_series, double *S_series);
/* Allocation-free measurement.  No counterpart in the reference: an entry
 * point for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  It
 * measures a lattice the caller holds on the current device as three FP32
 * SoA planes nx, ny, nz (site i + Nx*(j + Ny*k)) and writes
 * out[9] = {E, S, Qxx, Qxy, Qxz, Qyy, Qyz, Qzz, n_active} (device memory),
 * asynchronously on `stream` (a hipStream_t; NULL for the null stream).  For
 * the same lattice the values equal mc_measure's bit for bit (same kernels,
 * same launch shape).  Lattices: every site active, Nx a multiple of 4, x
 * and y periodic (even Ny); z periodic (even Nz, no surfaces) or free, with
 * no surfaces or with surfaces 0 and 1 on the planes k = 0 and k = Nz-1 (the
 * reference's slab, src/geometry_slab.cpp:47-67); the field strength is
 * params->elecE.  It calls no hipMalloc or hipFree: its partial sums live in
 * a fixed device array of MC_NOALLOC_MAX_PARTIALS rows of 8 doubles per GPU
 * (256 KiB, held until the process ends); a lattice needing more rows is
 * rejected with MC_EINVAL.  Calls that overlap on one GPU share that array
 * and race: order them. */
#define MC_NOALLOC_MAX_PARTIALS 4096
int mc_measure_noalloc(const mc_params *params, const float *nx, const float *ny, const float *nz, double *out,
                       void *stream);
/* FP32 site energies of every site (0 at pt==0 sites), float[Nt]. */
int mc_site_energies(mc_handle *h, float *out);

/* Lattice state transfer, AoS float[3*Nt] (src/evolve_thermal.cu:90,98). */
/* Local order parameter of every site for the director snapshot: S of the Q
 * tensor of the site and its active von Neumann neighbours, FP64 on the
 * device, rounded to float; NaN where pt == 0.  Replaces the host loop of
 * lattice_order_parameter over all sites (src/parameter_order.cpp:62-113,
 * called from print_n, src/io.cpp:50-80).  S_out: float[Nt]. */
int mc_local_order(mc_handle *h, float *S_out);
int mc_get_directors_aos(mc_handle *h, float *ni_aos);
int mc_set_directors_aos(mc_handle *h, const float *ni_aos);

/* Accepted count of the last completed sweep and the sweep counter. */
int mc_get_acceptance(mc_handle *h, uint32_t *accepted, uint32_t *sweeps_done);
/* Sets the sweep counter (the index of the next sweep, which keys the
 * generator).  With the directors (mc_get/set_directors_aos), ang_var
 * (mc_get/set_ang_var) and this counter a chain can be saved and resumed
 * exactly: the generator has no other state.  Extends the reference's
 * restart from a snapshot file (src/ic.cpp:88-119), which restores the
 * directors only. */
int mc_set_sweep_index(mc_handle *h, uint32_t sweep);

/* Parity hook: one pass over a site set with injected proposals (AoS
 * float[3*Nt]) and uniforms (float[Nt]) at the current T; set 0..7 is a
 * 2x2x2 sublattice tick, 8/9 a checkerboard colour.  accept/e_old/e_new are
 * float[Nt]/uint8[Nt] host arrays (any may be NULL) written at the sites of
 * the set.  Mirrors one callMCGPU launch (src/evolve.cu:93-203) with the CPU
 * accept rule u < exp(-dE/T) (src/evolve.cpp:160). */
int mc_pass(mc_handle *h, int set, const float *prop_aos, const float *u, uint8_t *accept, float *e_old,
            float *e_new);

/* One production sweep that also records, per site, the three uniforms, the
 * proposal, the accept flag and both energies (host arrays, any NULL), so the
 * CPU oracle can replay it pass by pass. */
int mc_record_sweep(mc_handle *h, float *u0, float *u1, float *u2, float *prop_aos, uint8_t *accept,
                    float *e_old, float *e_new);

/* ------------------------------------------------------------------------
 * Batched replicas.  nrep independent bulk lattices (periodic on every axis,
 * Nx % 4 == 0, even Ny and Nz, one potential for all) of any sizes, each with
 * its own parameters, temperature, seed/stream and proposal width, held in
 * one device allocation.  A sweep of all replicas is one slab launch over the
 * tiles of every replica (or two segment-kernel launches, one per colour),
 * then a one-block finisher, replayed as one captured graph; every replica
 * evolves bit-identically to the same lattice run alone with the segment
 * kernel.  No reference counterpart: the reference runs one
 * lattice per process (src/evolve.cu:265-270 keeps per-process statics). */
typedef struct mc_batch mc_batch;
int mc_batch_create(int nrep, const mc_params *params, const float *const *ni_aos, mc_batch **out);
int mc_batch_destroy(mc_batch *b);
int mc_batch_set_T(mc_batch *b, int rep, float T);
/* n sweeps of every replica, as graphs of at most 64 sweeps (as mc_sweeps_graph). */
int mc_batch_sweeps(mc_batch *b, int n);
int mc_batch_sync(mc_batch *b);
int mc_batch_time_sweeps(mc_batch *b, int n, float *ms);
/* Per-replica mean site energy and order parameter (double[nrep] each). */
int mc_batch_measure(mc_batch *b, double *E, double *S);
/* n measured sweeps; E/S are double[n*nrep], sweep-major. */
int mc_batch_measure_series(mc_batch *b, int n, double *E, double *S);
int mc_batch_get_directors_aos(mc_batch *b, int rep, float *ni_aos);
/* Per-replica accepted count of the last sweep, sweep counter, ang_var. */
int mc_batch_get_state(mc_batch *b, uint32_t *accepted, uint32_t *sweeps, float *ang_var);
/* Segment-kernel launch shape (grid 0 = default); selects the segment kernel. */
int mc_batch_set_tuning(mc_batch *b, int grid, int xcd);
/* Sweep kernel of the batch: MC_KERNEL_AUTO (the slab kernel where every
 * replica has a tile shape, else the segment kernel), MC_KERNEL_SLAB,
 * MC_KERNEL_SLAB_PERSISTENT or MC_KERNEL_X4.  All give every replica the
 * same chain.  MC_KERNEL_SLAB_PERSISTENT runs the slab tiles of up to 64
 * sweeps in one launch: as many workgroups as the device holds at once take
 * tiles from a per-sweep counter, with a grid-wide barrier between sweeps.
 * Its waits are bounded (mc_batch_set_barrier_limit): a barrier that does
 * not complete sets a fault that the next synchronising call (mc_batch_sync,
 * measurements, timing, state) reports as MC_EHIP. */
int mc_batch_set_kernel(mc_batch *b, int kind);
int mc_batch_get_kernel(mc_batch *b, int *kind);
/* Polls of a persistent launch's grid barrier before it gives up (default
 * 2^21, seconds of waiting; tests lower it to force the fault path). */
int mc_batch_set_barrier_limit(mc_batch *b, unsigned int polls);
/* Times segment grids and slab shapes on scratch copies and keeps the
 * fastest; grid = its workgroup count, xcd its tile schedule. */
int mc_batch_autotune(mc_batch *b, int *grid, int *xcd, float *best_us);

/* ------------------------------------------------------------------------
 * Learned proposal.  A 22-64-64-4 network (bf16 weights on the device, FP32
 * master copy) maps a site's six neighbour directors (sign-canonicalised), the
 * field vector elecE*e and 1/T to an axial von Mises-Fisher proposal
 *   q(n) = kappa cosh(kappa mu.n) / (4 pi sinh kappa),
 * mu = normalize(sum of canonical neighbours + o[0..2]),
 * log kappa = clamp(o[3], -6, 8).  Sweeps that use it accept a move o -> n with
 * probability min(1, exp(-dE/T) cosh(kappa mu.o) / cosh(kappa mu.n)), exact
 * Metropolis-Hastings because mu and kappa depend only on the neighbours.
 * Parameters are packed W1[64][22] b1[64] W2[64][64] b2[64] W3[4][64] b3[4]
 * (5892 floats).  No reference counterpart. */
#define MC_MLP_PARAMS 5892
typedef struct mc_mlp mc_mlp;
/* params NULL: deterministic initialisation from seed. */
int mc_mlp_create(const float *params, uint64_t seed, mc_mlp **out);
int mc_mlp_destroy(mc_mlp *m);
int mc_mlp_get_params(mc_mlp *m, float *params);
int mc_mlp_set_params(mc_mlp *m, const float *params);
/* Inference on host arrays: X[n][22] features, base[n][3] -> mu[n][3], logk[n]. */
int mc_mlp_infer(mc_mlp *m, int n, const float *X, const float *base, float *mu, float *logk);
/* Mean negative log-likelihood of dirs[n][3] and its gradient (float[5892]). */
int mc_mlp_grad(mc_mlp *m, int n, const float *X, const float *base, const float *dirs, double *loss, float *grad);
/* Appends n training neighbourhoods drawn at random sites of the lattice's
 * current state (its current T and field). */
int mc_mlp_dataset_add(mc_mlp *m, mc_handle *h, int n, uint32_t seed);
int mc_mlp_dataset_size(mc_mlp *m, int *n);
int mc_mlp_dataset_clear(mc_mlp *m);
/* Adam (beta 0.9/0.999) over the training set; mean loss per epoch. */
int mc_mlp_train(mc_mlp *m, int epochs, int batch, float lr, float *loss_per_epoch);
/* Attach (every > 0: every k-th sweep uses the network) or detach (NULL or 0). */
int mc_set_learned_proposal(mc_handle *h, mc_mlp *net, int every);
/* Learned sweeps follow each site's vMF move with a second Metropolis-
 * Hastings move, the reflection of the director through the network's mean
 * direction mu (a symmetric proposal; on by default).  on = 0 leaves the
 * vMF move alone. */
int mc_set_learned_reflection(mc_handle *h, int on);
/* One learned sweep recording per site mu (AoS), log kappa, the proposal
 * (AoS), the accept uniform, the decision and the FP32 energy change of the
 * vMF move, then the reflected director (AoS), its accept uniform, decision
 * and FP32 energy change (zeros when the reflection is off).  Any output may
 * be NULL. */
int mc_record_learned_sweep(mc_handle *h, float *mu, float *logk, float *prop_aos, float *u, uint8_t *accept,
                            float *dE, float *refl_aos, float *u2, uint8_t *accept2, float *dE2);

#ifdef __cplusplus
}
#endif
#endif
