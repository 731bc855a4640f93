// Host check of the equatorial-crossing root on the quartic dense output
// (z_root in gravlume_amd/csrc/trace_device.h), compiled with g++: on seeded
// DP5 steps whose z changes sign once, the binary32 root (bracket stopped at
// 2^-22, Newton quotient through the reciprocal) differs from the binary64
// root of the same step by no more than its binary32 condition number allows,
// 16 eps (|z0| + |h| sum |K_z|) / |z'(theta)| + 2^-21 (on a well-conditioned
// step, about 2e-6). A root stopped early, or a Newton update of the wrong
// sign, fails this bound.
// Prints the worst ratio to the bound; exits 1 on a violation.
#include <cmath>
#include <cstdio>
#include <random>

#include "../../gravlume_amd/csrc/ks_math.h"
#include "../../gravlume_amd/csrc/trace_device.h"

using namespace gl;

int main() {
  std::mt19937 rng(20261205);
  std::uniform_real_distribution<float> u(-1.0f, 1.0f);
  int cases = 0, bad = 0, well = 0;
  double worst = 0.0;
  for (int trial = 0; trial < 200000 && cases < 50000; ++trial) {
    float K[7][7] = {};
    // z' along the step: a common drift plus per-stage curvature terms
    const float drift = 2.0f * u(rng), curve = 0.3f * u(rng), jitter = 0.05f;
    const float c[7] = {0.0f, 0.2f, 0.3f, 0.8f, 8.0f / 9.0f, 1.0f, 1.0f};
    for (int i = 0; i < 7; ++i) K[i][3] = drift + curve * c[i] + jitter * u(rng);
    const float h = 0.05f + 2.0f * std::fabs(u(rng));
    const float z0 = 3.0f * u(rng);
    float x1, y1, z1;
    dense_position<float>(K, h, 0.0f, 0.0f, z0, 1.0f, x1, y1, z1);
    if (!((z0 < 0.0f && z1 > 0.0f) || (z0 > 0.0f && z1 < 0.0f))) continue;
    // one sign change on a fine grid: a single root in the step
    int changes = 0;
    float prev = z0;
    for (int k = 1; k <= 256; ++k) {
      float xs, ys, zs;
      dense_position<float>(K, h, 0.0f, 0.0f, z0, float(k) / 256.0f, xs, ys, zs);
      if ((zs < 0.0f) != (prev < 0.0f)) ++changes;
      prev = zs;
    }
    if (changes != 1) continue;
    ++cases;
    const float t32 = z_root<float, float>(K, h, z0, z1);
    const double t64 = z_root<double, float>(K, h, z0, z1);
    const double d = std::fabs(double(t32) - t64);
    double w[7];
    dense_slopes<double>(t64, w);
    double slope = 0.0, scale = std::fabs(double(z0));
    for (int i = 0; i < 7; ++i) {
      slope += w[i] * double(K[i][3]);
      scale += std::fabs(double(h) * double(K[i][3]));
    }
    slope = std::fabs(double(h) * slope);
    const double bound = 16.0 * 1.1920929e-7 * scale / slope + std::ldexp(1.0, -21);
    if (d / bound > worst) worst = d / bound;
    if (!(t32 >= 0.0f && t32 <= 1.0f) || d > bound) ++bad;
    if (slope >= 0.1 * scale) ++well;
  }
  std::printf("crossing roots: %d cases (%d well conditioned), worst |theta_f32 - theta_f64| / bound %.3g, violations %d\n",
              cases, well, worst, bad);
  return (cases >= 10000 && well >= 1000 && bad == 0) ? 0 : 1;
}
