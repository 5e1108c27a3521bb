// Learned proposal: a 22-64-64-4 MLP that maps a site's neighbourhood to an
// axial von Mises-Fisher proposal, trained and evaluated with bf16 MFMA
// (v_mfma_f32_16x16x32_bf16, FP32 accumulation) on gfx950, and used inside an
// exact Metropolis-Hastings sweep.  No reference counterpart.
//
// MFMA operand layout (16x16x32): lane l of the wave holds, for A (16 x 32,
// row-major M x K), row l%16 and the 8 K values 8*(l/16)..+7; for B (32 x 16,
// K x N), column l%16 and the same 8 K values; the 16x16 FP32 result D has
// D[4*(l/16)+i][l%16] in component i.  A and B always take the same K
// assignment, so a product is correct for any internal K order; only the
// documented M/N placement of D is relied on.
//
// Network and proposal, as in the CPU oracle:
//   x = features(neighbours, field, T) (22), h1 = relu(W1 x + b1),
//   h2 = relu(W2 h1 + b2), o = W3 h2 + b3,
//   mu = normalize(base + o[0..2]), log kappa = clamp(o[3], -6, 8),
//   q(n) = kappa cosh(kappa mu.n) / (4 pi sinh kappa).
// Because mu and kappa depend on the neighbours only, which a pass never
// changes at the updated site, the Metropolis-Hastings ratio of a move
// o -> n is exp(-dE/T) cosh(kappa mu.o) / cosh(kappa mu.n).
// A second move follows (unless switched off): the reflection of the
// site's director c through mu, r = 2 (mu.c) mu - c.  For a fixed mu it is
// an involution with unit Jacobian, a symmetric proposal, so its ratio is
// exp(-dE/T) alone.  Where mu is the principal axis of the site's local
// field (for Lebwohl-Lasher, of sum_b m_b m_b^T) the reflection conserves
// the energy exactly: an over-relaxation move, accepted almost always and
// far from the current director, driven by the network's mean direction.
#include "mc_kernels.h"
#include "mc_mlp.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int W1_OFF = 0, B1_OFF = 64 * 22, W2_OFF = B1_OFF + 64, B2_OFF = W2_OFF + 64 * 64,
              W3_OFF = B2_OFF + 64, B3_OFF = W3_OFF + 4 * 64;

__device__ __forceinline__ f4 mfma(bf16x8 a, bf16x8 b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x8 load8(const __bf16 *p) { return *reinterpret_cast<const bf16x8 *>(p); }

// B operand from a row-major [rows][ld] bf16 matrix: column n0 + l%16 of the
// product is row n0 + l%16 of W, K values k0 + 8*(l/16)..+7.
__device__ __forceinline__ bf16x8 bop(const __bf16 *W, int ld, int n0, int k0) {
  const int l = threadIdx.x & 63;
  return load8(W + (size_t)(n0 + (l & 15)) * ld + k0 + 8 * (l >> 4));
}

// Stores a D tile (+bias, optional ReLU) into an LDS row-major bf16 matrix.
__device__ __forceinline__ void store_tile(__bf16 *H, int ld, int m0, int n0, f4 c, const float *bias, bool relu) {
  const int l = threadIdx.x & 63;
  const int col = n0 + (l & 15);
  const float b = bias ? bias[col] : 0.0f;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    float v = c[i] + b;
    if (relu) v = v > 0.0f ? v : 0.0f;
    H[(m0 + 4 * (l >> 4) + i) * ld + col] = (__bf16)v;
  }
}

// Network weights in the operand-friendly bf16 layouts.
struct MLPDev {
  const __bf16 *W1b;   // [64][32]   (inputs 22..31 zero)
  const __bf16 *W2b;   // [64][64]
  const __bf16 *W3b;   // [16][64]   (rows 4..15 zero)
  const __bf16 *W3Tb;  // [64][32]   W3 transposed (cols 4..31 zero)
  const __bf16 *W2Tb;  // [64][64]   W2 transposed
  const float *b1, *b2, *b3;
};

// Forward of one 64-site block in transposed form: every layer computes
// D = W (units x K) * H^T (K x sites), weights as the A operand, sites as B.
// Lane (r, g) = (l % 16, l / 16) then holds units 16 mt + 4 g + i (i = 0..3)
// of site r of each 16-site tile, and the eight values it holds for tiles
// 2c and 2c+1 (units 32c + 4g + 0..3 and 32c + 16 + 4g + 0..3), after bias,
// ReLU and rounding to bf16, are exactly its B operand for the next layer's
// K chunk c, provided the weight operand uses the same K order (aop_hidden).
// So the hidden activations never leave registers; only the input rows
// (Xs, 64 x 32 bf16, row = site) and the four outputs per site (Os, 64 x 4
// FP32) go through LDS.  One wave.
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ bf16x8 aop_hidden(const __bf16 *W, int ld, int row, int c, int g) {
  const __bf16 *p = W + (size_t)row * ld + 32 * c + 4 * g;
  const bf16x4 lo = *reinterpret_cast<const bf16x4 *>(p);
  const bf16x4 hi = *reinterpret_cast<const bf16x4 *>(p + 16);
  return bf16x8{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
}

// ReLU, bf16 rounding and the D-tile to B-operand packing of K chunk c.
// Rounding first and then taking the ReLU as a signed 16-bit max with 0 on
// the bf16 bit patterns (v_pk_max_i16, two values per instruction) gives the
// same bits as rounding the ReLU: a negative input rounds to a value <= 0,
// whose pattern (sign bit set, or +0) maps to +0.
// (Conversions two at a time, one v_cvt_pk_bf16_f32 per pair of values,
// straight into the register pair layout of the operand.)
typedef short s16x8 __attribute__((ext_vector_type(8)));
typedef float f2v __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t cvt2(float x, float y) {
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(f2v{x, y}, bf16x2));
}
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ bf16x8 pack_hidden(const f4 &a, const f4 &b) {
  const u32x4 p = {cvt2(a[0], a[1]), cvt2(a[2], a[3]), cvt2(b[0], b[1]), cvt2(b[2], b[3])};
  const s16x8 w = __builtin_elementwise_max(__builtin_bit_cast(s16x8, p), s16x8{0, 0, 0, 0, 0, 0, 0, 0});
  return __builtin_bit_cast(bf16x8, w);
}

// The weight operands (rows 16 mt + l%16 of W1/W2, rows l%16 of W3; 14 KiB in
// all, L1-resident) are read where they are used rather than pinned in 56
// registers for the kernel's lifetime.
__device__ void forward64t(const MLPDev &N, const __bf16 *Xs, float *Os) {
  // Compiler-only barrier: keeps the weight loads here instead of hoisted
  // out of the caller's site loop into 56 long-lived registers.
  asm volatile("" ::: "memory");
  const int l = threadIdx.x & 63;
  const int r = l & 15, g = l >> 4;
#pragma unroll 1
  for (int t = 0; t < 4; ++t) {
    const bf16x8 x = load8(Xs + (16 * t + r) * 32 + 8 * g);
    f4 h1[4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
      h1[mt] = mfma(load8(N.W1b + (16 * mt + r) * 32 + 8 * g), x, *reinterpret_cast<const f4 *>(N.b1 + 16 * mt + 4 * g));
    const bf16x8 a0 = pack_hidden(h1[0], h1[1]), a1 = pack_hidden(h1[2], h1[3]);
    f4 h2[4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      f4 c = mfma(aop_hidden(N.W2b, 64, 16 * mt + r, 0, g), a0, *reinterpret_cast<const f4 *>(N.b2 + 16 * mt + 4 * g));
      h2[mt] = mfma(aop_hidden(N.W2b, 64, 16 * mt + r, 1, g), a1, c);
    }
    const bf16x8 b0 = pack_hidden(h2[0], h2[1]), b1 = pack_hidden(h2[2], h2[3]);
    f4 o = g == 0 ? *reinterpret_cast<const f4 *>(N.b3) : f4{0.0f, 0.0f, 0.0f, 0.0f};
    o = mfma(aop_hidden(N.W3b, 64, r, 0, g), b0, o);
    o = mfma(aop_hidden(N.W3b, 64, r, 1, g), b1, o);
    if (g == 0) *reinterpret_cast<f4 *>(Os + (16 * t + r) * 4) = o;  // outputs 0..3 of site 16t + r
  }
  __syncthreads();
}

// The same forward with the weight operands and biases in LDS (the block
// k_mlp_pack_ops writes, staged once per workgroup): operand b of lane l is
// the 16 bytes at Wl + 1024 b + 16 l, so every weight read is one
// conflict-free ds_read_b128 from a single per-lane address, and the biases
// are broadcast reads.  The input rows are 24 bf16 wide (Xs, 64 x 24): K
// values 24..31 are zero padding, so lanes with l/16 == 3 take a zero
// operand instead of reading them.  Xs and Os belong to the calling wave
// alone, which orders its own LDS writes and reads (wave_sync).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ void forward64l(const __bf16 *Wl, const float *Bl, const __bf16 *Xs, float *Os) {
  const int l = threadIdx.x & 63;
  const int r = l & 15, g = l >> 4;
  const bf16x8 *W = reinterpret_cast<const bf16x8 *>(Wl) + l;  // operand b at W[64 b]
#pragma unroll 1
  for (int t = 0; t < 4; t += 2) {
    // Two 16-site tiles per trip: their MFMA chains are independent, so the
    // second tile's products issue while the first tile's results are in
    // flight.  Compiler-only barrier: the operands are read again for every
    // trip instead of being hoisted out of the loop into long-lived registers.
    asm volatile("" ::: "memory");
    bf16x8 x[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) x[u] = g < 3 ? load8(Xs + (16 * (t + u) + r) * 24 + 8 * g) : bf16x8{};
    f4 h1[2][4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
      for (int u = 0; u < 2; ++u)
        h1[u][mt] = mfma(W[64 * mt], x[u], *reinterpret_cast<const f4 *>(Bl + 16 * mt + 4 * g));
    bf16x8 a0[2], a1[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      a0[u] = pack_hidden(h1[u][0], h1[u][1]);
      a1[u] = pack_hidden(h1[u][2], h1[u][3]);
    }
    f4 h2[2][4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        f4 c = mfma(W[64 * (4 + 2 * mt)], a0[u], *reinterpret_cast<const f4 *>(Bl + 64 + 16 * mt + 4 * g));
        h2[u][mt] = mfma(W[64 * (5 + 2 * mt)], a1[u], c);
      }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const bf16x8 b0 = pack_hidden(h2[u][0], h2[u][1]), b1 = pack_hidden(h2[u][2], h2[u][3]);
      f4 o = g == 0 ? *reinterpret_cast<const f4 *>(Bl + 128) : f4{0.0f, 0.0f, 0.0f, 0.0f};
      o = mfma(W[64 * 12], b0, o);
      o = mfma(W[64 * 13], b1, o);
      if (g == 0) *reinterpret_cast<f4 *>(Os + (16 * (t + u) + r) * 4) = o;
    }
  }
  wave_sync();
}

__device__ __forceinline__ void load_weights(const MLPDev &N, bf16x8 w1[4], bf16x8 w2[4][2], bf16x8 w3[2]) {
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    w1[nt] = bop(N.W1b, 32, 16 * nt, 0);
    w2[nt][0] = bop(N.W2b, 64, 16 * nt, 0);
    w2[nt][1] = bop(N.W2b, 64, 16 * nt, 32);
  }
  w3[0] = bop(N.W3b, 64, 0, 0);
  w3[1] = bop(N.W3b, 64, 0, 32);
}

// Features of one site, as orc_mlp_features.  The dot products that pick
// each neighbour's sign are fused multiply-adds here: only their signs are
// used, and those can differ from the unfused products only when d is a few
// ulps from zero, where either sign is a valid choice (a director and its
// negative are the same state).  Training samples and learned sweeps both
// take their features from this function.
__device__ __forceinline__ void features(const float m[6][3], float f0, float f1, float f2, float T, float x[22],
                                         float base[3]) {
  float c0 = m[0][0], c1 = m[0][1], c2 = m[0][2];
  const bool flip0 = c2 < 0.0f || (c2 == 0.0f && (c1 < 0.0f || (c1 == 0.0f && c0 < 0.0f)));
  if (flip0) {
    c0 = -c0;
    c1 = -c1;
    c2 = -c2;
  }
  x[0] = c0;
  x[1] = c1;
  x[2] = c2;
  base[0] = c0;
  base[1] = c1;
  base[2] = c2;
#pragma unroll
  for (int b = 1; b < 6; ++b) {
    const float d = fmaf(c2, m[b][2], fmaf(c1, m[b][1], c0 * m[b][0]));
    const float s = d < 0.0f ? -1.0f : 1.0f;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      x[3 * b + c] = s * m[b][c];
      base[c] = base[c] + s * m[b][c];
    }
  }
  x[18] = f0;
  x[19] = f1;
  x[20] = f2;
  x[21] = 1.0f / T;
}

__device__ __forceinline__ void write_xrow(__bf16 *Xs, int row, const float x[22], bool valid) {
#pragma unroll
  for (int c = 0; c < 4; ++c) {  // four 16-byte stores
    bf16x8 v;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int k = 8 * c + e;
      v[e] = (__bf16)((valid && k < 22) ? x[k] : 0.0f);
    }
    *reinterpret_cast<bf16x8 *>(Xs + row * 32 + 8 * c) = v;
  }
}

// Row `row` of the 24-wide input rows of forward64l (features 22, 23 zero).
__device__ __forceinline__ void write_xrow24(__bf16 *Xs, int row, const float x[22], bool valid) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    u32x4 v;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int k = 8 * c + 2 * e;
      const float x0 = (valid && k < 22) ? x[k] : 0.0f;
      const float x1 = (valid && k + 1 < 22) ? x[k + 1] : 0.0f;
      v[e] = cvt2(x0, x1);
    }
    *reinterpret_cast<u32x4 *>(Xs + row * 24 + 8 * c) = v;
  }
}

// Natural log of a normal positive float with the hardware log2
// (v_log_f32, within one ulp) and one multiply: two instructions where the
// library logf (with its denormal scaling and split ln 2) takes eleven.
__device__ __forceinline__ float hw_log(float x) { return __builtin_amdgcn_logf(x) * 0.693147181f; }

// log cosh(z) + log 2, FP32.
// log(1 + y) for y = e^{-2|z|} in (0, 1] with the hardware log: 1 + y is
// rounded once (absolute error <= 2^-24 in the result), where the library
// log1pf costs ~110 instructions per call.
__device__ __forceinline__ float lcosh2(float z) {
  const float a = fabsf(z);
  return a + hw_log(1.0f + __builtin_amdgcn_exp2f(a * -2.88539008f));  // e^{-2a} = 2^{-2a log2 e}
}

__device__ __forceinline__ void head(const float *o, const float base[3], float mu[3], float &logk) {
  float v0 = base[0] + o[0], v1 = base[1] + o[1], v2 = base[2] + o[2];
  float r2 = fmaf(v2, v2, fmaf(v1, v1, v0 * v0));
  // hardware reciprocal square root (1 ulp) on a normal r2; mu is a proposal
  // parameter, used alike in both directions of the move
  const float inv = r2 > 1e-30f ? __builtin_amdgcn_rsqf(r2) : 1.0f;
  mu[0] = v0 * inv;
  mu[1] = v1 * inv;
  mu[2] = v2 * inv;
  logk = __builtin_amdgcn_fmed3f(o[3], -6.0f, 8.0f);  // clamp to [-6, 8]
}

// ---------------------------------------------------------------------------

__global__ void k_mlp_pack(const float *__restrict__ P, __bf16 *W1b, __bf16 *W2b, __bf16 *W3b, __bf16 *W3Tb,
                           __bf16 *W2Tb, float *b1, float *b2, float *b3) {
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < 64 * 64; t += gridDim.x * blockDim.x) {
    const int j = t / 64, k = t % 64;
    W2b[t] = (__bf16)P[W2_OFF + t];
    W2Tb[k * 64 + j] = (__bf16)P[W2_OFF + t];
    if (k < 32) W1b[j * 32 + k] = (__bf16)(k < 22 ? P[W1_OFF + 22 * j + k] : 0.0f);
    if (j < 16) W3b[j * 64 + k] = (__bf16)(j < 4 ? P[W3_OFF + 64 * j + k] : 0.0f);
    if (j < 32) W3Tb[k * 32 + j] = (__bf16)(j < 4 ? P[W3_OFF + 64 * j + k] : 0.0f);
    if (t < 64) {
      b1[t] = P[B1_OFF + t];
      b2[t] = P[B2_OFF + t];
    }
    if (t < 4) b3[t] = P[B3_OFF + t];
  }
}

// The operand block of forward64l from the bf16 layouts: operands 0..3 are
// W1 rows 16 mt + l%16 (K 8(l/16)..+7), 4 + 2 mt + c the W2 chunks of
// aop_hidden, 12 + c the W3 chunks; then b1, b2, b3 and 12 zero floats.
__global__ void k_mlp_pack_ops(const __bf16 *__restrict__ W1b, const __bf16 *__restrict__ W2b,
                               const __bf16 *__restrict__ W3b, const float *__restrict__ b1,
                               const float *__restrict__ b2, const float *__restrict__ b3, void *ops) {
  __bf16 *W = reinterpret_cast<__bf16 *>(ops);
  float *B = reinterpret_cast<float *>(reinterpret_cast<char *>(ops) + 14 * 1024);
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < 14 * 64 + 144; t += gridDim.x * blockDim.x) {
    if (t >= 14 * 64) {
      const int k = t - 14 * 64;
      B[k] = k < 64 ? b1[k] : (k < 128 ? b2[k - 64] : (k < 132 ? b3[k - 128] : 0.0f));
      continue;
    }
    const int b = t >> 6, l = t & 63, r = l & 15, g = l >> 4;
    bf16x8 v;
    if (b < 4) {
      v = load8(W1b + (16 * b + r) * 32 + 8 * g);
    } else {
      const __bf16 *M = b < 12 ? W2b : W3b;
      const int row = b < 12 ? 16 * ((b - 4) >> 1) + r : r, c = b < 12 ? (b - 4) & 1 : b - 12;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        v[e] = M[row * 64 + 32 * c + 4 * g + e];
        v[4 + e] = M[row * 64 + 32 * c + 16 + 4 * g + e];
      }
    }
    *reinterpret_cast<bf16x8 *>(W + 8 * t) = v;
  }
}

// Inference over n neighbourhoods given as features X[n][22] and base[n][3].
__global__ __launch_bounds__(64) void k_mlp_infer(MLPDev N, const float *__restrict__ X,
                                                  const float *__restrict__ base, int n, float *__restrict__ mu,
                                                  float *__restrict__ logk) {
  __shared__ __align__(16) __bf16 Xs[64 * 32];
  __shared__ __align__(16) float Os[64 * 4];

  const int l = threadIdx.x;
  for (int blk = blockIdx.x; blk * 64 < n; blk += gridDim.x) {
    const int s = blk * 64 + l;
    const bool valid = s < n;
    float x[22];
#pragma unroll
    for (int k = 0; k < 22; ++k) x[k] = valid ? X[(size_t)s * 22 + k] : 0.0f;
    write_xrow(Xs, l, x, valid);
    __syncthreads();
    forward64t(N, Xs, Os);
    if (valid) {
      const float b[3] = {base[3 * (size_t)s], base[3 * (size_t)s + 1], base[3 * (size_t)s + 2]};
      float m[3], lk;
      head(Os + l * 4, b, m, lk);
      mu[3 * (size_t)s] = m[0];
      mu[3 * (size_t)s + 1] = m[1];
      mu[3 * (size_t)s + 2] = m[2];
      logk[s] = lk;
    }
    __syncthreads();
  }
}

// Gradient of the mean negative log-likelihood over samples [0, n) (each
// wave takes 32-sample batches): forward, loss head, backward with bf16 MFMA
// products and FP32 accumulation; per-wave gradient tiles are added to
// grad[5892] (FP32 atomics) and the loss sum to loss[0].
__global__ __launch_bounds__(64) void k_mlp_grad(MLPDev N, const float *__restrict__ X,
                                                 const float *__restrict__ base, const float *__restrict__ dirs, int n,
                                                 float *__restrict__ grad, double *__restrict__ loss) {
  __shared__ __align__(16) __bf16 Xs[32 * 32];
  __shared__ __align__(16) __bf16 H1s[32 * 64];
  __shared__ __align__(16) __bf16 H2s[32 * 64];
  __shared__ __align__(16) __bf16 dOs[32 * 32];
  __shared__ __align__(16) __bf16 dH2s[32 * 64];
  __shared__ __align__(16) __bf16 dH1s[32 * 64];
  __shared__ float Os[32 * 16];
  const int l = threadIdx.x;
  const int r = l & 15, g = l >> 4;
  bf16x8 w1[4], w2[4][2], w3[2], w3t[4], w2t[4][2];
  load_weights(N, w1, w2, w3);
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    w3t[nt] = bop(N.W3Tb, 32, 16 * nt, 0);
    w2t[nt][0] = bop(N.W2Tb, 64, 16 * nt, 0);
    w2t[nt][1] = bop(N.W2Tb, 64, 16 * nt, 32);
  }
  f4 gW1[4][2], gW2[4][4], gW3[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    gW3[a] = f4{0, 0, 0, 0};
    for (int b = 0; b < 2; ++b) gW1[a][b] = f4{0, 0, 0, 0};
    for (int b = 0; b < 4; ++b) gW2[a][b] = f4{0, 0, 0, 0};
  }
  float gb1[4] = {0, 0, 0, 0}, gb2[4] = {0, 0, 0, 0}, gb3 = 0.0f;
  double lsum = 0.0;
  for (int blk = blockIdx.x; blk * 32 < n; blk += gridDim.x) {
    // inputs: lanes 0..31 one sample each
    if (l < 32) {
      const int s = blk * 32 + l;
      float x[22];
#pragma unroll
      for (int k = 0; k < 22; ++k) x[k] = s < n ? X[(size_t)s * 22 + k] : 0.0f;
      write_xrow(Xs, l, x, s < n);
    }
    __syncthreads();
    // forward (2 M-tiles of 16)
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
      const bf16x8 a = load8(Xs + (16 * mt + r) * 32 + 8 * g);
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) store_tile(H1s, 64, 16 * mt, 16 * nt, mfma(a, w1[nt], f4{0, 0, 0, 0}), N.b1, true);
    }
    __syncthreads();
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
      const bf16x8 a0 = load8(H1s + (16 * mt + r) * 64 + 8 * g);
      const bf16x8 a1 = load8(H1s + (16 * mt + r) * 64 + 32 + 8 * g);
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
        f4 c = mfma(a0, w2[nt][0], f4{0, 0, 0, 0});
        c = mfma(a1, w2[nt][1], c);
        store_tile(H2s, 64, 16 * mt, 16 * nt, c, N.b2, true);
      }
    }
    __syncthreads();
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
      const bf16x8 a0 = load8(H2s + (16 * mt + r) * 64 + 8 * g);
      const bf16x8 a1 = load8(H2s + (16 * mt + r) * 64 + 32 + 8 * g);
      f4 c = mfma(a0, w3[0], f4{0, 0, 0, 0});
      c = mfma(a1, w3[1], c);
#pragma unroll
      for (int i = 0; i < 4; ++i) Os[(16 * mt + 4 * g + i) * 16 + r] = c[i] + (r < 4 ? N.b3[r] : 0.0f);
    }
    __syncthreads();
    // loss head: lanes 0..31, dOut into dOs (bf16, columns 0..3)
    if (l < 32) {
      const int s = blk * 32 + l;
      float d[4] = {0, 0, 0, 0};
      if (s < n) {
        const float *o = Os + l * 16;
        const float b0 = base[3 * (size_t)s], b1v = base[3 * (size_t)s + 1], b2v = base[3 * (size_t)s + 2];
        const float v0 = b0 + o[0], v1 = b1v + o[1], v2 = b2v + o[2];
        float rr = sqrtf(v0 * v0 + v1 * v1 + v2 * v2);
        if (rr == 0.0f) rr = 1.0f;
        const float m0 = v0 / rr, m1 = v1 / rr, m2 = v2 / rr;
        float lk = o[3];
        const bool clamped = lk < -6.0f || lk > 8.0f;
        lk = fminf(fmaxf(lk, -6.0f), 8.0f);
        const float k = __expf(lk);
        const float n0 = dirs[3 * (size_t)s], n1 = dirs[3 * (size_t)s + 1], n2 = dirs[3 * (size_t)s + 2];
        const float c = m0 * n0 + m1 * n1 + m2 * n2;
        const float z = k * c;
        // NLL = -log k + log 4pi + log sinh k - log cosh z
        const float lsinh = k + log1pf(-__expf(-2.0f * k)) - 0.69314718f;
        lsum += (double)(-lk + 2.5310242f + lsinh - (lcosh2(z) - 0.69314718f));
        const float th = tanhf(z);
        d[3] = clamped ? 0.0f : (-1.0f + k / tanhf(k) - z * th);
        const float dm0 = -k * th * n0, dm1 = -k * th * n1, dm2 = -k * th * n2;
        const float pr = dm0 * m0 + dm1 * m1 + dm2 * m2;
        d[0] = (dm0 - pr * m0) / rr;
        d[1] = (dm1 - pr * m1) / rr;
        d[2] = (dm2 - pr * m2) / rr;
      }
#pragma unroll
      for (int c = 0; c < 32; ++c) dOs[l * 32 + c] = (__bf16)(c < 4 ? d[c] : 0.0f);
      // bias of the output layer: column sums of dOut (lane-private, reduced below)
      Os[l * 16 + 4] = d[0];
      Os[l * 16 + 5] = d[1];
      Os[l * 16 + 6] = d[2];
      Os[l * 16 + 7] = d[3];
    }
    __syncthreads();
    if (l < 4)
      for (int s = 0; s < 32; ++s) gb3 += Os[s * 16 + 4 + l];
    // dW3[o][k] += sum_s dOut[s][o] H2[s][k]: A = dOut^T (o x s), B = H2 (s x k)
    {
      bf16x8 a;
#pragma unroll
      for (int e = 0; e < 8; ++e) a[e] = dOs[(8 * g + e) * 32 + r];
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
        bf16x8 b;
#pragma unroll
        for (int e = 0; e < 8; ++e) b[e] = H2s[(8 * g + e) * 64 + 16 * nt + r];
        gW3[nt] = mfma(a, b, gW3[nt]);
      }
    }
    // dH2 = (dOut W3) * relu'(h2)
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
      const bf16x8 a = load8(dOs + (16 * mt + r) * 32 + 8 * g);
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
        const f4 c = mfma(a, w3t[nt], f4{0, 0, 0, 0});
        const int col = 16 * nt + r;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int row = 16 * mt + 4 * g + i;
          const float v = (float)H2s[row * 64 + col] > 0.0f ? c[i] : 0.0f;
          gb2[nt] += v;
          dH2s[row * 64 + col] = (__bf16)v;
        }
      }
    }
    __syncthreads();
    // dW2[j][k] += sum_s dH2[s][j] H1[s][k]
#pragma unroll
    for (int jt = 0; jt < 4; ++jt) {
      bf16x8 a;
#pragma unroll
      for (int e = 0; e < 8; ++e) a[e] = dH2s[(8 * g + e) * 64 + 16 * jt + r];
#pragma unroll
      for (int kt = 0; kt < 4; ++kt) {
        bf16x8 b;
#pragma unroll
        for (int e = 0; e < 8; ++e) b[e] = H1s[(8 * g + e) * 64 + 16 * kt + r];
        gW2[jt][kt] = mfma(a, b, gW2[jt][kt]);
      }
    }
    // dH1 = (dH2 W2) * relu'(h1)
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
      const bf16x8 a0 = load8(dH2s + (16 * mt + r) * 64 + 8 * g);
      const bf16x8 a1 = load8(dH2s + (16 * mt + r) * 64 + 32 + 8 * g);
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
        f4 c = mfma(a0, w2t[nt][0], f4{0, 0, 0, 0});
        c = mfma(a1, w2t[nt][1], c);
        const int col = 16 * nt + r;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int row = 16 * mt + 4 * g + i;
          const float v = (float)H1s[row * 64 + col] > 0.0f ? c[i] : 0.0f;
          gb1[nt] += v;
          dH1s[row * 64 + col] = (__bf16)v;
        }
      }
    }
    __syncthreads();
    // dW1[j][k] += sum_s dH1[s][j] X[s][k]
#pragma unroll
    for (int jt = 0; jt < 4; ++jt) {
      bf16x8 a;
#pragma unroll
      for (int e = 0; e < 8; ++e) a[e] = dH1s[(8 * g + e) * 64 + 16 * jt + r];
#pragma unroll
      for (int kt = 0; kt < 2; ++kt) {
        bf16x8 b;
#pragma unroll
        for (int e = 0; e < 8; ++e) b[e] = Xs[(8 * g + e) * 32 + 16 * kt + r];
        gW1[jt][kt] = mfma(a, b, gW1[jt][kt]);
      }
    }
    __syncthreads();
  }
  // flush: D tiles hold rows 4g+i (output index within the tile), column r
#pragma unroll
  for (int jt = 0; jt < 4; ++jt) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int j = 16 * jt + 4 * g + i;
      for (int kt = 0; kt < 2; ++kt) {
        const int k = 16 * kt + r;
        if (k < 22) atomicAdd(&grad[W1_OFF + 22 * j + k], gW1[jt][kt][i]);
      }
      for (int kt = 0; kt < 4; ++kt) atomicAdd(&grad[W2_OFF + 64 * j + 16 * kt + r], gW2[jt][kt][i]);
    }
  }
#pragma unroll
  for (int nt = 0; nt < 4; ++nt)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int o = 4 * g + i;
      if (o < 4) atomicAdd(&grad[W3_OFF + 64 * o + 16 * nt + r], gW3[nt][i]);
    }
  // bias sums: lanes r, r+16, r+32, r+48 hold partial sums of column r
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    float v1 = gb1[nt], v2 = gb2[nt];
    v1 += __shfl_xor(v1, 16, 64);
    v1 += __shfl_xor(v1, 32, 64);
    v2 += __shfl_xor(v2, 16, 64);
    v2 += __shfl_xor(v2, 32, 64);
    if (l < 16) {
      atomicAdd(&grad[B1_OFF + 16 * nt + l], v1);
      atomicAdd(&grad[B2_OFF + 16 * nt + l], v2);
    }
  }
  if (l < 4) atomicAdd(&grad[B3_OFF + l], gb3);
  for (int off = 32; off > 0; off >>= 1) lsum += __shfl_down(lsum, off, 64);
  if (l == 0) atomicAdd(loss, lsum);
}

// Adam on the FP32 master copy with the mean gradient grad/nsamp; clears grad.
__global__ void k_mlp_adam(float *__restrict__ P, float *__restrict__ grad, float *__restrict__ m1,
                           float *__restrict__ m2, int np, float inv_n, float lr, float beta1, float beta2, float eps,
                           float bc1, float bc2) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < np; p += gridDim.x * blockDim.x) {
    const float g = grad[p] * inv_n;
    const float a = beta1 * m1[p] + (1.0f - beta1) * g;
    const float b = beta2 * m2[p] + (1.0f - beta2) * g * g;
    m1[p] = a;
    m2[p] = b;
    P[p] -= lr * (a / bc1) / (sqrtf(b / bc2) + eps);
    grad[p] = 0.0f;
  }
}

__global__ void k_fill_f32(float *p, int n, float v) {
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) p[t] = v;
}

// Gathers the six neighbours of site (i,j,k) of an all-active lattice with
// the lattice's boundary types (bulk neighbourhoods only).
// PER: every axis periodic, so every neighbour exists and the six indices
// are the site's own index plus or minus a stride, wrapped by one select.
template <bool PER = false>
__device__ __forceinline__ void gather6(const MCDevParams &P, const float *nx, const float *ny, const float *nz,
                                        int i, int j, int k, float m[6][3], bool on[6]) {
  const int X = P.nx, Y = P.ny, Z = P.nz;
  if (PER) {
    const int XY = X * Y, s = i + X * (j + Y * k);
    const int idx[6] = {i + 1 == X ? s + 1 - X : s + 1,     i == 0 ? s - 1 + X : s - 1,
                        j + 1 == Y ? s + X - XY : s + X,    j == 0 ? s - X + XY : s - X,
                        k + 1 == Z ? s + XY - P.nt : s + XY, k == 0 ? s - XY + P.nt : s - XY};
    // 32-bit byte offsets from a uniform base (no 64-bit address per load)
#pragma unroll
    for (int b = 0; b < 6; ++b) {
      on[b] = true;
      const uint32_t off = (uint32_t)idx[b] * 4u;
      m[b][0] = *reinterpret_cast<const float *>(reinterpret_cast<const char *>(nx) + off);
      m[b][1] = *reinterpret_cast<const float *>(reinterpret_cast<const char *>(ny) + off);
      m[b][2] = *reinterpret_cast<const float *>(reinterpret_cast<const char *>(nz) + off);
    }
    return;
  }
  int ip = i + 1, im = i - 1, jp = j + 1, jm = j - 1, kp = k + 1, km = k - 1;
  on[0] = on[1] = on[2] = on[3] = on[4] = on[5] = true;
  if (ip == X) { if (P.per[0]) ip = 0; else on[0] = false; }
  if (im < 0) { if (P.per[0]) im = X - 1; else on[1] = false; }
  if (jp == Y) { if (P.per[1]) jp = 0; else on[2] = false; }
  if (jm < 0) { if (P.per[1]) jm = Y - 1; else on[3] = false; }
  if (kp == Z) { if (P.per[2]) kp = 0; else on[4] = false; }
  if (km < 0) { if (P.per[2]) km = Z - 1; else on[5] = false; }
  const int idx[6] = {ip + X * (j + Y * k), im + X * (j + Y * k), i + X * (jp + Y * k),
                      i + X * (jm + Y * k), i + X * (j + Y * kp), i + X * (j + Y * km)};
#pragma unroll
  for (int b = 0; b < 6; ++b) {
    const int s = on[b] ? idx[b] : 0;
    m[b][0] = on[b] ? nx[s] : 0.0f;
    m[b][1] = on[b] ? ny[s] : 0.0f;
    m[b][2] = on[b] ? nz[s] : 0.0f;
  }
}

// Training samples from the current lattice: random sites (Philox keyed by
// seed, counter = sample index) -> features, base, centre director.
__global__ void k_mlp_dataset(MCDevParams P, const float *__restrict__ nx, const float *__restrict__ ny,
                              const float *__restrict__ nz, const MCDevState *st, int n, uint32_t seed,
                              float *__restrict__ X, float *__restrict__ base, float *__restrict__ dirs) {
  const float T = st->T, E = st->elecE;
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
    const MCU4 r = mcd_philox(MCU4{(uint32_t)s, seed, P.stream, 0x4D53u}, P.seed_lo, P.seed_hi);
    const int site = (int)(((uint64_t)r.x * (uint64_t)P.nt) >> 32);
    const int i = site % P.nx, j = (site / P.nx) % P.ny, k = site / (P.nx * P.ny);
    float m[6][3];
    bool on[6];
    gather6(P, nx, ny, nz, i, j, k, m, on);
    float x[22], b[3];
    features(m, E * P.eX, E * P.eY, E * P.eZ, T, x, b);
#pragma unroll
    for (int c = 0; c < 22; ++c) X[(size_t)s * 22 + c] = x[c];
    base[3 * (size_t)s] = b[0];
    base[3 * (size_t)s + 1] = b[1];
    base[3 * (size_t)s + 2] = b[2];
    dirs[3 * (size_t)s] = nx[site];
    dirs[3 * (size_t)s + 1] = ny[site];
    dirs[3 * (size_t)s + 2] = nz[site];
  }
}

// Draws from the axial vMF mixture (Wood 1994 for the polar cosine on S^2,
// w = 1 + log(v + (1 - v) e^{-2 kappa}) / kappa with v in (0, 1], a uniform
// azimuth around mu, then the sign of the mixture component).
__device__ __forceinline__ void sample_axial_vmf(const float mu[3], float kappa, float u_w, float u_phi, bool neg,
                                                 float q[3]) {
  const float v = 1.0f - u_w;  // (0, 1]
  // (hardware reciprocal and square root, 1 ulp: the draw is a sample, and
  // the Metropolis-Hastings ratio uses the density, not these operations)
  // (the log's argument is at least v >= 2^-24, a normal float)
  float w = fmaf(hw_log(fmaf(u_w, __builtin_amdgcn_exp2f(kappa * -2.88539008f), v)),  // 1 - v = u_w exactly
                 __builtin_amdgcn_rcpf(kappa), 1.0f);
  w = fminf(fmaxf(w, -1.0f), 1.0f);
  const float st = __builtin_amdgcn_sqrtf(fmaxf(0.0f, fmaf(-w, w, 1.0f)));
  // orthonormal e, f perpendicular to the unit mu, branch free (Duff et al.,
  // "Building an orthonormal basis, revisited", JCGT 6(1), 2017)
  const float sgn = mu[2] >= 0.0f ? 1.0f : -1.0f;
  const float ia = -__builtin_amdgcn_rcpf(sgn + mu[2]);
  const float bb = mu[0] * mu[1] * ia;
  const float si = sgn * ia;
  const float e0 = fmaf(si * mu[0], mu[0], 1.0f), e1 = sgn * bb, e2 = -sgn * mu[0];
  const float f0 = bb, f1 = fmaf(mu[1] * mu[1], ia, sgn), f2 = -mu[1];
  const float ph = u_phi;  // revolutions
  const float cs = __builtin_amdgcn_cosf(ph), sn = __builtin_amdgcn_sinf(ph);
  float q0 = fmaf(st, fmaf(cs, e0, sn * f0), w * mu[0]);
  float q1 = fmaf(st, fmaf(cs, e1, sn * f1), w * mu[1]);
  float q2 = fmaf(st, fmaf(cs, e2, sn * f2), w * mu[2]);
  // |q|^2 = 1 + d with d a few ulps: q (1.5 - 0.5 |q|^2) is unit to second order
  const float iq = fmaf(-0.5f, fmaf(q2, q2, fmaf(q1, q1, q0 * q0)), 1.5f);
  const float sg = neg ? -iq : iq;
  q[0] = q0 * sg;
  q[1] = q1 * sg;
  q[2] = q2 * sg;
}

// One colour pass of the learned-proposal Metropolis-Hastings sweep over an
// all-active lattice.  A workgroup of kLearnWaves waves stages the network
// (weight operands and biases, 14.6 KiB) into LDS once; then each wave takes
// 64-site blocks on its own (blocks wave, wave + kLearnWaves * gridDim.x,
// ...) with its own input rows and outputs in LDS, so the only workgroup
// barriers are after staging and at the final count.
constexpr int kLearnWaves = 4;

template <int POT, bool RECORD, bool PER>
__global__ __launch_bounds__(64 * kLearnWaves) __attribute__((amdgpu_waves_per_eu(4))) void k_learned_pass(
    MCDevParams P, const uint4 *__restrict__ ops, float *__restrict__ nx, float *__restrict__ ny,
    float *__restrict__ nz, MCDevState *st, int colour, int finalize, int reflect, MCLearnRecord rec) {
  __shared__ __align__(16) uint4 Wst[MC_MLP_OPS_BYTES / 16];
  __shared__ __align__(16) __bf16 Xall[kLearnWaves][64 * 24];
  __shared__ __align__(16) float Oall[kLearnWaves][64 * 4];
  __shared__ uint32_t wsum[kLearnWaves];
  for (int i = threadIdx.x; i < MC_MLP_OPS_BYTES / 16; i += 64 * kLearnWaves) Wst[i] = ops[i];
  __syncthreads();
  const __bf16 *Wl = reinterpret_cast<const __bf16 *>(Wst);
  const float *Bl = reinterpret_cast<const float *>(Wst + 14 * 64);
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  __bf16 *Xs = Xall[wave];
  float *Os = Oall[wave];

  const float T = st->T, E = st->elecE;
  const float invT = 1.0f / T;
  const uint32_t sweep = st->sweep;
  const double cA = (double)(-P.A) * 1.5;
  const float cAf = (float)cA;
  const int l = threadIdx.x & 63;
  const int total = mcd_set_slots(P, 8 + colour);
  // Periodic lattice with Nx/2 a multiple of 64: a block is 64 consecutive
  // colour sites of one row, so its row (j, k) is one scalar for the wave.
  const bool rowblk = PER && ((P.nx >> 1) & 63) == 0;
  uint32_t acc = 0;
  for (int blk = blockIdx.x * kLearnWaves + wave; blk * 64 < total; blk += gridDim.x * kLearnWaves) {
    const int t = blk * 64 + l;
    bool valid;
    int s;
    float m[6][3];
    bool on[6];
    float b[3];
    if (rowblk) {
      // (its own gather, so the row's offsets stay scalar)
      const int t0 = blk * 64;
      const int row = (int)mcd_div((uint32_t)t0, P.div_hx);
      const int k = (int)mcd_div((uint32_t)row, P.div_ny);
      const int j = row - k * P.ny;
      const int i = 2 * (t0 - row * (int)P.div_hx.d + l) + ((j + k + colour) & 1);
      valid = true;  // every slot of a block is a site: total is a multiple of 64
      s = i + P.nx * (j + P.ny * k);
      gather6<PER>(P, nx, ny, nz, i, j, k, m, on);
    } else {
      int i = 0, j = 0, k = 0;
      valid = t < total && mcd_set_site(P, 8 + colour, t, i, j, k);
      if (!valid) i = j = k = 0;  // a slot past an odd free Nx: gather in-bounds, result unused
      s = valid ? i + P.nx * (j + P.ny * k) : 0;
      gather6<PER>(P, nx, ny, nz, i, j, k, m, on);
    }
    {
      float x[22];
      features(m, E * P.eX, E * P.eY, E * P.eZ, T, x, b);
      // (a lane past the end holds site 0's features: the forward pass keeps
      // rows apart, and that lane's move is never taken)
      write_xrow24(Xs, l, x, true);
    }
    wave_sync();
    forward64l(Wl, Bl, Xs, Os);
    float mu[3], lk;
    head(Os + l * 4, b, mu, lk);
    const float kappa = __expf(lk);
    const MCU4 r = mcd_philox(MCU4{(uint32_t)s, sweep, P.stream, 0x4D4Cu}, P.seed_lo, P.seed_hi);
    float q[3];
    sample_axial_vmf(mu, kappa, mcd_u01(r.x), mcd_u01(r.y), (r.z & 1u) != 0u, q);
    const float o0 = valid ? nx[s] : 1.0f, o1 = valid ? ny[s] : 0.0f, o2 = valid ? nz[s] : 0.0f;
    float dE;
    float dot_o[6], dot_q[6];  // director . neighbour for the old and the proposed director (LL)
    if (POT == MCD_LL) {
      float a2 = 0.0f;
#pragma unroll
      for (int bb = 0; bb < 6; ++bb) {
        dot_o[bb] = fmaf(o2, m[bb][2], fmaf(o1, m[bb][1], o0 * m[bb][0]));
        dot_q[bb] = fmaf(q[2], m[bb][2], fmaf(q[1], m[bb][1], q[0] * m[bb][0]));
        if (!on[bb]) continue;
        a2 = fmaf(dot_q[bb] - dot_o[bb], dot_q[bb] + dot_o[bb], a2);
      }
      dE = cAf * a2;
    } else {
      double so = 0.0, sn = 0.0;
#pragma unroll
      for (int bb = 0; bb < 6; ++bb) {
        if (!on[bb]) continue;
        so += (double)mcd_bond<POT>(P, cA, bb, o0, o1, o2, m[bb][0], m[bb][1], m[bb][2]);
        sn += (double)mcd_bond<POT>(P, cA, bb, q[0], q[1], q[2], m[bb][0], m[bb][1], m[bb][2]);
      }
      dE = (float)sn - (float)so;
    }
    if (P.elecA != 0.0f) dE += mcd_field(P, E, q[0], q[1], q[2]) - mcd_field(P, E, o0, o1, o2);
    const float co = fmaf(mu[2], o2, fmaf(mu[1], o1, mu[0] * o0));
    const float cn = fmaf(mu[2], q[2], fmaf(mu[1], q[1], mu[0] * q[0]));
    const float lr = lcosh2(kappa * co) - lcosh2(kappa * cn);
    const float xlog = fmaf(-dE, invT, lr);
    const float u = mcd_u01(r.w);
    const bool a = valid && u < __expf(fminf(xlog, 80.0f));
    // The reflection move from the director c the first move left.  Its
    // accept uniform is the top 24 bits of word z (whose bit 0 chose the
    // vMF mixture sign).  The reflected vector is renormalised to first
    // order (|r| = 1 up to a few ulps already).
    float c0 = a ? q[0] : o0, c1 = a ? q[1] : o1, c2 = a ? q[2] : o2;
    float r0 = c0, r1 = c1, r2 = c2, dE2 = 0.0f, u2 = 0.0f;
    bool a2 = false;
    if (reflect) {
      const float dm2 = 2.0f * fmaf(mu[2], c2, fmaf(mu[1], c1, mu[0] * c0));
      r0 = fmaf(dm2, mu[0], -c0);
      r1 = fmaf(dm2, mu[1], -c1);
      r2 = fmaf(dm2, mu[2], -c2);
      const float rn = fmaf(-0.5f, fmaf(r2, r2, fmaf(r1, r1, r0 * r0)), 1.5f);
      r0 *= rn;
      r1 *= rn;
      r2 *= rn;
      if (POT == MCD_LL) {
        float b2 = 0.0f;
#pragma unroll
        for (int bb = 0; bb < 6; ++bb) {
          if (!on[bb]) continue;
          const float ac = a ? dot_q[bb] : dot_o[bb];  // c . m, computed for the first move
          const float ar = fmaf(r2, m[bb][2], fmaf(r1, m[bb][1], r0 * m[bb][0]));
          b2 = fmaf(ar - ac, ar + ac, b2);
        }
        dE2 = cAf * b2;
      } else {
        double so = 0.0, sn = 0.0;
#pragma unroll
        for (int bb = 0; bb < 6; ++bb) {
          if (!on[bb]) continue;
          so += (double)mcd_bond<POT>(P, cA, bb, c0, c1, c2, m[bb][0], m[bb][1], m[bb][2]);
          sn += (double)mcd_bond<POT>(P, cA, bb, r0, r1, r2, m[bb][0], m[bb][1], m[bb][2]);
        }
        dE2 = (float)sn - (float)so;
      }
      if (P.elecA != 0.0f) dE2 += mcd_field(P, E, r0, r1, r2) - mcd_field(P, E, c0, c1, c2);
      u2 = mcd_u01(r.z);
      a2 = valid && u2 < __expf(fminf(-dE2 * invT, 80.0f));
    }
    if (a2) {
      c0 = r0;
      c1 = r1;
      c2 = r2;
    }
    if (a || a2) {
      nx[s] = c0;
      ny[s] = c1;
      nz[s] = c2;
    }
    acc += a ? 1u : 0u;
    if (RECORD && valid) {
      rec.mu[3 * (size_t)s] = mu[0];
      rec.mu[3 * (size_t)s + 1] = mu[1];
      rec.mu[3 * (size_t)s + 2] = mu[2];
      rec.logk[s] = lk;
      rec.prop[3 * (size_t)s] = q[0];
      rec.prop[3 * (size_t)s + 1] = q[1];
      rec.prop[3 * (size_t)s + 2] = q[2];
      rec.u[s] = u;
      rec.acc[s] = a ? 1 : 0;
      rec.dE[s] = dE;
      if (rec.refl && reflect) {
        rec.refl[3 * (size_t)s] = r0;
        rec.refl[3 * (size_t)s + 1] = r1;
        rec.refl[3 * (size_t)s + 2] = r2;
        rec.u2[s] = u2;
        rec.acc2[s] = a2 ? 1 : 0;
        rec.dE2[s] = dE2;
      }
    }
    wave_sync();  // this wave's Os reads before the next block's Xs/Os writes
  }
  // acceptance bookkeeping; ang_var (unused by the learned proposal) is left
  // as it is for plain sweeps interleaved with learned ones
  uint32_t v = acc;
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if (l == 0) wsum[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    v = 0;
#pragma unroll
    for (int w = 0; w < kLearnWaves; ++w) v += wsum[w];
  }
  mcd_finish_sweep(P, st, v, finalize != 0, /*adapt=*/false);
}

MLPDev dev_of(const MCMLPBuffers &B) {
  MLPDev N;
  N.W1b = B.W1b;
  N.W2b = B.W2b;
  N.W3b = B.W3b;
  N.W3Tb = B.W3Tb;
  N.W2Tb = B.W2Tb;
  N.b1 = B.b1;
  N.b2 = B.b2;
  N.b3 = B.b3;
  return N;
}

}  // namespace

void mc_mlp_launch_pack(const MCMLPBuffers &B, hipStream_t s) {
  k_mlp_pack<<<16, 256, 0, s>>>(B.params, B.W1b, B.W2b, B.W3b, B.W3Tb, B.W2Tb, B.b1, B.b2, B.b3);
  k_mlp_pack_ops<<<4, 256, 0, s>>>(B.W1b, B.W2b, B.W3b, B.b1, B.b2, B.b3, B.ops);
}

void mc_mlp_launch_infer(const MCMLPBuffers &B, hipStream_t s, const float *X, const float *base, int n, float *mu,
                         float *logk) {
  const int grid = (n + 63) / 64 < 4096 ? (n + 63) / 64 : 4096;
  k_mlp_infer<<<grid > 0 ? grid : 1, 64, 0, s>>>(dev_of(B), X, base, n, mu, logk);
}

void mc_mlp_launch_grad(const MCMLPBuffers &B, hipStream_t s, const float *X, const float *base, const float *dirs,
                        int n, int grid) {
  k_mlp_grad<<<grid, 64, 0, s>>>(dev_of(B), X, base, dirs, n, B.grad, B.loss);
}

void mc_mlp_launch_adam(const MCMLPBuffers &B, hipStream_t s, int nsamp, float lr, int step) {
  const float b1 = 0.9f, b2 = 0.999f;
  const float bc1 = 1.0f - powf(b1, (float)step), bc2 = 1.0f - powf(b2, (float)step);
  k_mlp_adam<<<24, 256, 0, s>>>(B.params, B.grad, B.m1, B.m2, MC_MLP_NPARAM, 1.0f / nsamp, lr, b1, b2, 1e-8f, bc1,
                                bc2);
}

void mc_mlp_launch_fill(float *p, int n, float v, hipStream_t s) { k_fill_f32<<<24, 256, 0, s>>>(p, n, v); }

void mc_mlp_launch_dataset(const MCLaunch &L, int n, uint32_t seed, float *X, float *base, float *dirs) {
  const int grid = (n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096;
  k_mlp_dataset<<<grid, 256, 0, L.stream>>>(L.P, L.nx, L.ny, L.nz, L.st, n, seed, X, base, dirs);
}

// Workgroups of the pass's kernel that one CU holds at once (registers and
// LDS), asked of the runtime once per instantiation.
template <int POT, bool PER>
int learned_resident() {
  static int n = 0;
  if (n == 0) {
    int r = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&r, reinterpret_cast<const void *>(&k_learned_pass<POT, false, PER>),
                                                     64 * kLearnWaves, 0) != hipSuccess || r < 1) {
      (void)hipGetLastError();
      r = 4;
    }
    n = r;
  }
  return n;
}

int mc_mlp_learned_grid(const MCDevParams &P, int slots, int cu_count) {
  const bool per = P.per[0] && P.per[1] && P.per[2];
  int res;
  switch (P.potential) {
    case MCD_GHRL:
      res = per ? learned_resident<MCD_GHRL, true>() : learned_resident<MCD_GHRL, false>();
      break;
    case MCD_PEAR:
      res = per ? learned_resident<MCD_PEAR, true>() : learned_resident<MCD_PEAR, false>();
      break;
    default:
      res = per ? learned_resident<MCD_LL, true>() : learned_resident<MCD_LL, false>();
      break;
  }
  const int blocks = (slots + 63) / 64;
  int grid = (blocks + kLearnWaves - 1) / kLearnWaves;
  if (grid > cu_count * res) grid = cu_count * res;
  return grid > 0 ? grid : 1;
}

void mc_mlp_launch_learned_pass(const MCLaunch &L, const MCMLPBuffers &B, int colour, int grid, int finalize,
                                int reflect, const MCLearnRecord *rec) {
  const MCLearnRecord none{};
  const MCLearnRecord &R = rec ? *rec : none;
  const uint4 *ops = reinterpret_cast<const uint4 *>(B.ops);
  const int threads = 64 * kLearnWaves;
  const bool per = L.P.per[0] && L.P.per[1] && L.P.per[2];
#define MC_LEARNED_(POT_, REC_, PER_) \
  k_learned_pass<POT_, REC_, PER_><<<grid, threads, 0, L.stream>>>(L.P, ops, L.nx, L.ny, L.nz, L.st, colour, finalize, reflect, R)
#define MC_LEARNED(POT_)                  \
  if (rec) {                              \
    if (per) MC_LEARNED_(POT_, true, true);   \
    else MC_LEARNED_(POT_, true, false);      \
  } else {                                \
    if (per) MC_LEARNED_(POT_, false, true);  \
    else MC_LEARNED_(POT_, false, false);     \
  }
  switch (L.P.potential) {
    case MCD_GHRL:
      MC_LEARNED(MCD_GHRL);
      break;
    case MCD_PEAR:
      MC_LEARNED(MCD_PEAR);
      break;
    default:
      MC_LEARNED(MCD_LL);
      break;
  }
#undef MC_LEARNED
#undef MC_LEARNED_
}
