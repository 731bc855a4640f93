// Host check of the interval capture certificate (gravlume_amd/csrc/certificate.h)
// against the f64 oracle, compiled together with the oracle sources.
//
// For each scene, the initial ray of every k-th pixel is built in f64 by the
// oracle, rounded to f32 as the device holds it, and its constants of motion
// and future radial rate evaluated with the device templates (ks_math.h).
//   * soundness: every certified ray must end at the horizon in the oracle's
//     f64 trace (strict-v1 policy);
//   * the reference's default Kerr scene (a = 0.8, r = 30 M, 60 deg, FOV 45
//     deg, 1280x720): the certified fraction is printed for comparison with
//     the reference's 56,937 / 921,600 px (6.178%).
// Prints "cert <scene> sampled certified violations"; exits non-zero on any
// violation.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <vector>

#include "../../gravlume_amd/csrc/certificate.h"
#include "../../gravlume_amd/csrc/ks_math.h"
#include "../../oracle/oracle.hpp"

using namespace gl;

struct SceneSpec {
  const char* name;
  double a, q, r_obs, incl_deg, fov_deg;
  uint32_t w, h, stride;
};

static int run(const SceneSpec& S, long* certified_out, long* sampled_out) {
  // outgoing chart: backward rays cross the (past) horizon regularly
  const double chart = -1.0;
  const double th = S.incl_deg * M_PI / 180.0;
  const double cs = chart * S.a;
  double ev[4] = {0.0, S.r_obs * std::sin(th), cs * std::sin(th), S.r_obs * std::cos(th)};
  double target[4] = {0, 0, 0, 0}, up[3] = {0, 0, 1};
  orc::Observation obs;
  if (orc::build_observation(1.0, S.a, S.q, chart, ev, target, up, 1.0, S.w, S.h,
                             S.fov_deg * M_PI / 180.0, &obs))
    return -1;
  orc::Policy pol = orc::strict_v1();
  orc::Events evc;
  evc.has_escape = 1;
  evc.escape_radius = 1000.0;
  evc.horizon_offset = 1e-6;
  const float a_f = float(S.a), q_f = float(S.q);
  const float rp = float(1.0 + std::sqrt(1.0 - S.a * S.a - S.q * S.q));
  std::atomic<long> certified{0}, sampled{0}, violations{0};
  std::atomic<uint32_t> next_row{0};
  auto worker = [&]() {
    for (;;) {
      const uint32_t py = next_row.fetch_add(S.stride);
      if (py >= S.h) return;
      for (uint32_t px = 0; px < S.w; px += S.stride) {
        double sx, sy;
        if (orc::sight_plane(obs, px, py, 0.5, 0.5, &sx, &sy)) continue;
        orc::S8 y;
        double sight[4], fn, nres;
        if (orc::initial_ray_sight(obs, sx, sy, &y, sight, &fn, &nres)) continue;
        ++sampled;
        const float x = float(y[1]), yy = float(y[2]), z = float(y[3]);
        const float E = float(-y[4]), pxv = float(y[5]), pyv = float(y[6]), pzv = float(y[7]);
        Geom<float> g = geometry<float>(x, yy, z, 1.0f, a_f, a_f * a_f, q_f * q_f, float(chart));
        Deriv<float> d = rhs_from_geometry<float>(g, x, yy, z, pxv, pyv, pzv, E, a_f, float(chart));
        const float dr = radial_rate<float>(g, d);
        const float L = x * pyv - yy * pxv;
        const float Q = carter<float>(g, x, yy, z, pxv, pyv, pzv, E, a_f * a_f);
        if (!capture_certified(1.0f, a_f, q_f, rp, g.r, x * x + yy * yy, dr, E, L, Q)) continue;
        ++certified;
        orc::Outcome out;
        orc::trace(obs.st, pol, evc, y, -1.0, &out);
        if (out.termination != orc::kTermHorizon) {
          if (++violations <= 10)
            std::printf("VIOLATION %s pixel (%u, %u): oracle class %d\n", S.name, px, py,
                        out.termination);
        }
      }
    }
  };
  std::vector<std::thread> pool;
  for (int t = 0; t < 8; ++t) pool.emplace_back(worker);
  for (auto& t : pool) t.join();
  std::printf("cert %s sampled %ld certified %ld (%.3f%%) violations %ld\n", S.name,
              sampled.load(), certified.load(), 100.0 * certified.load() / std::max(1L, sampled.load()),
              violations.load());
  *certified_out = certified;
  *sampled_out = sampled;
  return int(violations.load());
}

int main() {
  const SceneSpec scenes[] = {
      {"default-kerr", 0.8, 0.0, 30.0, 60.0, 45.0, 1280, 720, 4},
      {"kn-near", 0.5, 0.3, 12.0, 75.0, 60.0, 256, 256, 4},
      {"schwarzschild", 0.0, 0.0, 20.0, 30.0, 50.0, 256, 256, 4},
      {"kerr-0.9", 0.9, 0.0, 40.0, 85.0, 30.0, 256, 256, 4},
  };
  int bad = 0;
  long total_cert = 0;
  for (const auto& s : scenes) {
    long c = 0, n = 0;
    int v = run(s, &c, &n);
    if (v != 0) bad = 1;
    total_cert += c;
  }
  if (total_cert == 0) bad = 1;
  return bad;
}
