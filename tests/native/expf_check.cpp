// Compares the engine's expf (mc_expf.h, the function the GPU's exact accept
// path evaluates) with the host's glibc expf over every FP32 input.  Prints
// the number of inputs checked, the number of differing results and the
// first differing input, and exits 0 only when none differ.  NaN results
// match when both are NaN.
#define MCX_HD
#include "../../mclics-backend_amd/csrc/mc_expf.h"

#include <math.h>
#include <stdio.h>
#include <string.h>

int main() {
  long long bad = 0;
  unsigned first = 0xffffffffu;
#pragma omp parallel for schedule(static) reduction(+ : bad)
  for (long long b = 0; b < (1ll << 32); ++b) {
    const unsigned ub = (unsigned)b;
    float x;
    memcpy(&x, &ub, 4);
    const float g = expf(x), m = mcx_expf(x);
    unsigned gb, mb;
    memcpy(&gb, &g, 4);
    memcpy(&mb, &m, 4);
    if (gb != mb && !(g != g && m != m)) {
      ++bad;
#pragma omp critical
      if (ub < first) first = ub;
    }
  }
  float fx;
  memcpy(&fx, &first, 4);
  printf("checked %lld mismatches %lld first %08x (%.9g)\n", 1ll << 32, bad, first, bad ? fx : 0.0f);
  return bad == 0 ? 0 : 1;
}
