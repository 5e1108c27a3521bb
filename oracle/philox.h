/*
 * Philox4x32-10 counter-based generator, C restatement for the oracle.
 *
 * TEST INFRASTRUCTURE: part of the CPU oracle. Only tests, the smoke check and
 * the bench's CPU-baseline leg use it.
 *
 * Written from Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as
 * 1, 2, 3", SC'11, section 4.2 (Philox-4x32: two 32x32->64 multiplies per
 * round, Weyl key schedule).  The multipliers and Weyl increments are the
 * published constants of that paper.  The reference itself draws from GSL
 * ranlxs0 (src/evolve.cpp:136,137,160) and cuRAND XORWOW
 * (src/evolve.cu:25-31,143,145,193); neither stream is reproducible here, so
 * the GPU path and this oracle share this generator instead.
 */
#ifndef MC_ORACLE_PHILOX_H_
#define MC_ORACLE_PHILOX_H_
#include <stdint.h>

#define MC_PHILOX_M0 0xD2511F53u
#define MC_PHILOX_M1 0xCD9E8D57u
#define MC_PHILOX_W0 0x9E3779B9u
#define MC_PHILOX_W1 0xBB67AE85u

typedef struct { uint32_t v[4]; } mc_u32x4;

static inline void mc_philox_round(uint32_t c[4], const uint32_t k[2]) {
  uint64_t p0 = (uint64_t)MC_PHILOX_M0 * c[0];
  uint64_t p1 = (uint64_t)MC_PHILOX_M1 * c[2];
  uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
  uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
  uint32_t n0 = hi1 ^ c[1] ^ k[0];
  uint32_t n2 = hi0 ^ c[3] ^ k[1];
  c[0] = n0;
  c[1] = lo1;
  c[2] = n2;
  c[3] = lo0;
}

/* Ten rounds, key bumped between rounds. */
static inline mc_u32x4 mc_philox4x32_10(mc_u32x4 ctr, uint32_t key0, uint32_t key1) {
  uint32_t k[2] = {key0, key1};
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k[0] += MC_PHILOX_W0;
      k[1] += MC_PHILOX_W1;
    }
    mc_philox_round(ctr.v, k);
  }
  return ctr;
}

/* 24-bit uniform on [0,1): exactly representable in float. */
static inline float mc_u01_24(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }

#endif
