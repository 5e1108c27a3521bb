// Learned-proposal network: device buffers and launchers (mc_mlp.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mc_kernels.h"

#define MC_MLP_NPARAM (64 * 22 + 64 + 64 * 64 + 64 + 4 * 64 + 4)

struct MCMLPBuffers {
  float *params;  // FP32 master copy [5892]
  float *grad;    // FP32 gradient sums [5892]
  float *m1, *m2; // Adam moments
  double *loss;   // loss sum
  __bf16 *W1b, *W2b, *W3b, *W3Tb, *W2Tb;  // bf16 operand layouts
  float *b1, *b2, *b3;
  // The forward pass's 14 weight operands lane by lane (bf16, 1 KiB each),
  // then b1, b2, b3 (FP32, 144 floats): the block the learned pass stages
  // into LDS with plain 16-byte copies.
  void *ops;
};

#define MC_MLP_OPS_BYTES (14 * 1024 + 144 * 4)

// Per-site capture of one learned pass (device arrays of Nt; all null = off).
struct MCLearnRecord {
  float *mu;    // [3*Nt]
  float *logk;  // [Nt]
  float *prop;  // [3*Nt]
  float *u;     // [Nt]
  uint8_t *acc; // [Nt]
  float *dE;    // [Nt]
  // second move (reflection through mu), recorded when refl is non-null
  float *refl;   // [3*Nt] reflected director
  float *u2;     // [Nt] its accept uniform
  uint8_t *acc2; // [Nt]
  float *dE2;    // [Nt] its FP32 energy change
};

void mc_mlp_launch_pack(const MCMLPBuffers &B, hipStream_t s);
void mc_mlp_launch_infer(const MCMLPBuffers &B, hipStream_t s, const float *X, const float *base, int n, float *mu,
                         float *logk);
void mc_mlp_launch_grad(const MCMLPBuffers &B, hipStream_t s, const float *X, const float *base, const float *dirs,
                        int n, int grid);
void mc_mlp_launch_adam(const MCMLPBuffers &B, hipStream_t s, int nsamp, float lr, int step);
void mc_mlp_launch_fill(float *p, int n, float v, hipStream_t s);
void mc_mlp_launch_dataset(const MCLaunch &L, int n, uint32_t seed, float *X, float *base, float *dirs);
// Workgroups of a learned pass over `slots` sites: one 64-site block per
// wave, up to the workgroups every CU holds at once.
int mc_mlp_learned_grid(const MCDevParams &P, int slots, int cu_count);
// reflect: follow each vMF move with the reflection through mu (mc_mlp.hip).
void mc_mlp_launch_learned_pass(const MCLaunch &L, const MCMLPBuffers &B, int colour, int grid, int finalize,
                                int reflect, const MCLearnRecord *rec);

struct mc_mlp;
struct mc_handle;
const MCMLPBuffers &mc_mlp_buffers(const mc_mlp *m);
int mc_mlp_device(const mc_mlp *m);
