// Host check of the GHRL screen's error bound (mclics-backend_amd/csrc/
// mc_ghrl_screen.h): over random moves, neighbourhoods and elastic
// constants, the screen's FP32 dE must lie within 920 u M of the dE of the
// reference's own arithmetic (the oracle's bonds summed in double and
// rounded per site, src/geometry.cpp:19-72), u = 2^-24, M the bond bound;
// moves it flags as near a chiral sign change are counted, not checked.
// Prints the largest distance seen in units of u M.  Exit status 1 on any
// violation.  Built with g++ -ffp-contract=off together with the oracle.
#define MCX_HD
#include "../../mclics-backend_amd/csrc/mc_ghrl_screen.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include <random>

extern "C" {
#include "../../oracle/mc_oracle.h"
}

static void unit(std::mt19937_64 &g, float v[3]) {
  std::normal_distribution<double> N(0.0, 1.0);
  double a = N(g), b = N(g), c = N(g), r = sqrt(a * a + b * b + c * c);
  v[0] = (float)(a / r);
  v[1] = (float)(b / r);
  v[2] = (float)(c / r);
}

int main(int argc, char **argv) {
  const long trials = argc > 1 ? atol(argv[1]) : 2000000;
  std::mt19937_64 g(12345);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  const float ks[][4] = {{1.0f, 0.7f, 1.3f, 0.0f}, {1.0f, 0.7f, 1.3f, 16.0f}, {1.48f, 1.0f, 2.28f, 0.0f},
                         {1.0f, 1.0f, 1.0f, 8.0f}, {0.6f, 1.4f, 0.9f, -5.0f}};
  const float r[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
  double worst = 0.0;
  long tiny_n = 0, bad = 0, checked = 0;
  for (long t = 0; t < trials; ++t) {
    const float *k = ks[t % 5];
    orc_params p = {};
    p.potential = ORC_GHRL;
    orc_set_ghrl(&p, k[0], k[1], k[2], k[3]);
    const float M = 2.0f * (fabsf(p.ghrl_rho) + fabsf(p.ghrl_lambda)) + 1.12f * fabsf(p.ghrl_mu) +
                    fabsf(p.ghrl_nu) + fabsf(p.ghrl_sigma);
    float o[3], q[3], m[6][3];
    unit(g, o);
    const int mode = (int)(t / 5 % 4);
    if (mode == 0) {
      unit(g, q);  // a fresh direction
    } else {
      // a small rotation, as the sweep proposes
      float d[3];
      unit(g, d);
      const double eps = mode == 1 ? 0.3 : (mode == 2 ? 1e-3 : 1e-6);
      double v[3] = {o[0] + eps * d[0], o[1] + eps * d[1], o[2] + eps * d[2]};
      double n = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
      for (int c = 0; c < 3; ++c) q[c] = (float)(v[c] / n);
    }
    for (int b = 0; b < 6; ++b) {
      if (U(g) < 0.5) {
        // an ordered neighbour near o
        float d[3];
        unit(g, d);
        double v[3] = {o[0] + 0.2 * d[0], o[1] + 0.2 * d[1], o[2] + 0.2 * d[2]};
        double n = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        for (int c = 0; c < 3; ++c) m[b][c] = (float)(v[c] / n);
      } else {
        unit(g, m[b]);
      }
    }
    const uint32_t on = (t % 7 == 0) ? 0x0Fu : (t % 11 == 0 ? 0x1Fu : 0x3Fu);  // free z faces drop bonds
    double so = 0.0, sn = 0.0;
    for (int b = 0; b < 6; ++b) {
      if (!((on >> b) & 1u)) continue;
      so += (double)orc_bond(&p, o, m[b], r[b]);
      sn += (double)orc_bond(&p, q, m[b], r[b]);
    }
    const float ref = (float)sn - (float)so;
    const McGhrlConst c = mcx_ghrl_const(p.ghrl_rho, p.ghrl_lambda, p.ghrl_mu, p.ghrl_nu, p.ghrl_sigma);
    bool tiny = false;
    const float scr = mcx_ghrl_delta(c, o, q, m, on, &tiny);
    if (tiny) {
      ++tiny_n;
      continue;
    }
    ++checked;
    const double dist = fabs((double)scr - (double)ref) / (ldexp(1.0, -24) * M);
    if (dist > worst) worst = dist;
    if (dist > 920.0) {
      if (bad < 5) printf("violation: k=(%g,%g,%g,%g) dist %.1f u M scr %.9g ref %.9g\n", k[0], k[1], k[2], k[3], dist,
                          scr, ref);
      ++bad;
    }
  }
  printf("checked %ld tiny %ld worst %.2f u M violations %ld\n", checked, tiny_n, worst, bad);
  return bad ? 1 : 0;
}
