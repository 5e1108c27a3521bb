// expf as the reference's CPU accept test evaluates it.
//
// The reference decides a move with u < exp(-dE/T) on a float argument
// (src/evolve.cpp:160), which libstdc++ resolves to std::exp(float), i.e.
// glibc's expf.  glibc 2.35's expf is not correctly rounded: it is the
// table-driven method of ARM's optimized-routines (Szabolcs Nagy, 2017;
// glibc sysdeps/ieee754/flt-32/e_expf.c with e_exp2f_data.c), and on x86-64
// CPUs with FMA and AVX2 the dynamic loader selects the build of that file
// compiled with -mfma, in which the compiler fused the multiply-adds.  The
// method: x*32/ln2 = k + r with k rounded to the nearest integer by adding
// 1.5*2^52, |r| <= 1/2; 2^(k/32) = 2^(k>>5) * T[k & 31]; 2^(r/32) ~ 1 + C2 r
// + C1 r^2 + C0 r^3; everything in FP64 and rounded once to FP32.
//
// This is that evaluation written out with explicit fused multiply-adds
// where that build has them, so it gives the same bits on the host and on
// the GPU (v_fma_f64 is an exact IEEE fma).  Over every one of the 2^32
// FP32 inputs it equals glibc 2.35's expf on an FMA-capable x86-64 host; the
// build of the same file without FMA differs from it at two inputs.
//
// Table: T[i] = bits of 2^(i/32) rounded to the nearest double, minus
// i << 47, so that adding k << 47 to T[k & 31] also scales by 2^(k >> 5).
// Polynomial: the published minimax coefficients of 2^(r/32) on
// |r| <= 1/2 (powers of 32 folded in, which is exact).
#pragma once
#include <stdint.h>

#ifndef MCX_HD
#define MCX_HD __host__ __device__
#endif

MCX_HD inline uint64_t mcx_f64_bits(double d) {
  uint64_t u;
  __builtin_memcpy(&u, &d, 8);
  return u;
}
MCX_HD inline double mcx_f64_from(uint64_t u) {
  double d;
  __builtin_memcpy(&d, &u, 8);
  return d;
}

#define MCX_EXPF_TABLE {\
      0x3ff0000000000000ull, 0x3fefd9b0d3158574ull, 0x3fefb5586cf9890full, 0x3fef9301d0125b51ull,\
      0x3fef72b83c7d517bull, 0x3fef54873168b9aaull, 0x3fef387a6e756238ull, 0x3fef1e9df51fdee1ull,\
      0x3fef06fe0a31b715ull, 0x3feef1a7373aa9cbull, 0x3feedea64c123422ull, 0x3feece086061892dull,\
      0x3feebfdad5362a27ull, 0x3feeb42b569d4f82ull, 0x3feeab07dd485429ull, 0x3feea47eb03a5585ull,\
      0x3feea09e667f3bcdull, 0x3fee9f75e8ec5f74ull, 0x3feea11473eb0187ull, 0x3feea589994cce13ull,\
      0x3feeace5422aa0dbull, 0x3feeb737b0cdc5e5ull, 0x3feec49182a3f090ull, 0x3feed503b23e255dull,\
      0x3feee89f995ad3adull, 0x3feeff76f2fb5e47ull, 0x3fef199bdd85529cull, 0x3fef3720dcef9069ull,\
      0x3fef5818dcfba487ull, 0x3fef7c97337b9b5full, 0x3fefa4afa2a490daull, 0x3fefd0765b6e4540ull}

// Entry i of the table (for copying it into fast memory).
MCX_HD inline uint64_t mcx_expf_table_entry(int i) {
  constexpr uint64_t kTab[32] = MCX_EXPF_TABLE;
  return kTab[i];
}

// expf with the table read from tab (host memory, device global memory or
// LDS: the caller picks where the 32 entries live).
MCX_HD inline float mcx_expf_with(float x, const uint64_t *tab) {
  constexpr double kInvLn2N = 0x1.71547652b82fep+0 * 32;  // 32 / ln 2
  constexpr double kShift = 0x1.8p+52;                    // rounds to an integer when added
  constexpr double C0 = 0x1.c6af84b912394p-5 / 32 / 32 / 32;
  constexpr double C1 = 0x1.ebfce50fac4f3p-3 / 32 / 32;
  constexpr double C2 = 0x1.62e42ff0c52d6p-1 / 32;
  uint32_t ux;
  __builtin_memcpy(&ux, &x, 4);
  const uint32_t abstop = (ux >> 20) & 0x7ffu;
  if (abstop >= 0x42bu) {                       // |x| >= 88 or NaN
    if (ux == 0xff800000u) return 0.0f;         // -inf
    if (abstop >= 0x7f8u) return x + x;         // +inf, NaN
    if (x > 0x1.62e42ep6f) return __builtin_inff();  // overflow
    if (x < -0x1.9fe368p6f) return 0.0f;        // underflow
  }
  const double xd = (double)x;
  double kd = __builtin_fma(kInvLn2N, xd, kShift);
  const uint64_t ki = mcx_f64_bits(kd);
  kd -= kShift;
  const double r = __builtin_fma(kInvLn2N, xd, -kd);
  const double s = mcx_f64_from(tab[ki & 31u] + (ki << 47));
  const double z = __builtin_fma(C0, r, C1);
  const double r2 = r * r;
  double y = __builtin_fma(C2, r, 1.0);
  y = __builtin_fma(z, r2, y);
  y = y * s;
  return (float)y;
}

MCX_HD inline float mcx_expf(float x) {
  constexpr uint64_t kTab[32] = MCX_EXPF_TABLE;
  return mcx_expf_with(x, kTab);
}
