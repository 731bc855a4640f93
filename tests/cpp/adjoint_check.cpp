// Host checks of the discrete adjoint (gravlume_amd/csrc/adjoint.h), in double:
//  1. rhs_vjp against central differences of the forward RHS, for every
//     input, at seeded points on both branches of the radius root;
//  2. rhs_only against geometry + rhs_from_geometry (same derivative);
//  3. the frozen-step trajectory adjoint (step_adjoint over N DP5 steps, then
//     terminal_reverse through the dense-output equator crossing with its
//     implicit fraction theta*) against central differences of the same map
//     y0 -> y(theta*) with respect to the initial state, E, a and Q^2.
// Prints the worst relative errors; exits non-zero above 1e-7 / 1e-6.
#include <cmath>
#include <cstdio>
#include <random>

#include "../../gravlume_amd/csrc/adjoint.h"

using namespace gl;

static void rhs(const double v[10], double s, double out[6]) {
  // v = x y z px py pz E a Q2 M
  Geom<double> g = geometry<double>(v[0], v[1], v[2], v[9], v[7], v[7] * v[7], v[8], s);
  Deriv<double> d = rhs_from_geometry<double>(g, v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7], s);
  out[0] = d.dx; out[1] = d.dy; out[2] = d.dz; out[3] = d.dpx; out[4] = d.dpy; out[5] = d.dpz;
}

struct Params {
  double y0[6], E, a, q2;
};

static SceneT<double> scene_of(const Params& P) {
  SceneT<double> sc{};
  sc.mass = 1.0;
  sc.spin = P.a;
  sc.spin2 = P.a * P.a;
  sc.charge2 = P.q2;
  sc.chart = -1.0;
  return sc;
}

// Frozen-step map: N steps of size h, then the terminal step of size hT to
// the equator crossing on its dense output.
static void trajectory(const Params& P, int N, double h, double hT, double yT[6]) {
  SceneT<double> sc = scene_of(P);
  double y[6];
  for (int c = 0; c < 6; ++c) y[c] = P.y0[c];
  for (int n = 0; n < N; ++n) advance<double>(sc, y, P.E, h);
  double K[7][6], Kt[7][7], dt;
  double th = terminal_stages<double>(sc, y, P.E, hT, K, Kt);
  dense_state<double>(Kt, hT, y, th, yT, dt);
}

static double trajectory_check() {
  Params P = {{30.0, 2.0, 8.0, -0.95, 0.12, -0.32}, 1.0, 0.9, 0.04};
  // find the step that crosses the equator with h = 1
  const double h = 1.0;
  SceneT<double> sc = scene_of(P);
  double y[6];
  for (int c = 0; c < 6; ++c) y[c] = P.y0[c];
  int N = 0;
  for (;; ++N) {
    double y1[6];
    for (int c = 0; c < 6; ++c) y1[c] = y[c];
    advance<double>(sc, y1, P.E, h);
    if (y1[2] <= 0.0) break;
    for (int c = 0; c < 6; ++c) y[c] = y1[c];
    if (N > 200) return 1.0;
  }
  const double hT = h;
  const double lT[6] = {0.3, -0.7, 0.2, 1.1, -0.4, 0.9};
  // adjoint
  double lam[6];
  RhsBar<double> acc{};
  {
    double K[7][6], Kt[7][7];
    double th = terminal_stages<double>(sc, y, P.E, hT, K, Kt);
    if (!terminal_reverse<double>(sc, y, P.E, hT, K, th, lT, lam, acc)) return 1.0;
    for (int n = N - 1; n >= 0; --n) {
      double yn[6];
      for (int c = 0; c < 6; ++c) yn[c] = P.y0[c];
      for (int m = 0; m < n; ++m) advance<double>(sc, yn, P.E, h);
      step_adjoint<double>(sc, yn, P.E, h, lam, acc);
    }
  }
  const double got[9] = {lam[0], lam[1], lam[2], lam[3], lam[4], lam[5], acc.E,
                         acc.spin + 2.0 * P.a * acc.spin2, acc.charge2};
  double worst = 0.0;
  for (int i = 0; i < 9; ++i) {
    Params Pp = P, Pm = P;
    double* vp = i < 6 ? &Pp.y0[i] : (i == 6 ? &Pp.E : (i == 7 ? &Pp.a : &Pp.q2));
    double* vm = i < 6 ? &Pm.y0[i] : (i == 6 ? &Pm.E : (i == 7 ? &Pm.a : &Pm.q2));
    const double eps = 1e-6 * std::fmax(1.0, std::fabs(*vp));
    *vp += eps;
    *vm -= eps;
    double fp[6], fm[6];
    trajectory(Pp, N, h, hT, fp);
    trajectory(Pm, N, h, hT, fm);
    double fd = 0.0;
    for (int c = 0; c < 6; ++c) fd += lT[c] * (fp[c] - fm[c]) / (2.0 * eps);
    const double err = std::fabs(fd - got[i]) / std::fmax(1e-3, std::fabs(fd));
    if (err > 1e-6) std::printf("trajectory input %d: fd %.12g adjoint %.12g\n", i, fd, got[i]);
    worst = std::fmax(worst, err);
  }
  std::printf("trajectory: %d frozen steps + terminal crossing\n", N);
  return worst;
}

int main() {
  std::mt19937_64 rng(20261203);
  std::uniform_real_distribution<double> U(-1.0, 1.0);
  double worst = 0.0;
  int cases = 0;
  for (int t = 0; t < 400; ++t) {
    double v[10];
    const double s = (t & 1) ? 1.0 : -1.0;
    const double scale = (t % 4 < 2) ? 12.0 : 0.6;  // far points, and points inside the ring
    v[0] = scale * U(rng); v[1] = scale * U(rng); v[2] = scale * U(rng);
    if (t % 8 == 7) { v[0] *= 1e-3; v[1] *= 1e-3; }  // near the axis
    v[3] = U(rng); v[4] = U(rng); v[5] = U(rng);
    v[6] = 1.0 + 0.2 * U(rng);
    v[7] = 0.95 * U(rng);
    v[8] = 0.3 * (U(rng) + 1.0) * 0.5;
    v[9] = 1.0;
    double r2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    if (r2 < 0.05) continue;
    double lam[6];
    for (double& l : lam) l = U(rng);
    RhsBar<double> b{};
    rhs_vjp<double>(v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[9], v[7], v[7] * v[7], v[8], s, lam, &b);
    const double got[9] = {b.x, b.y, b.z, b.px, b.py, b.pz, b.E, b.spin + 2.0 * v[7] * b.spin2, b.charge2};
    for (int i = 0; i < 9; ++i) {
      double h = 1e-6 * std::fmax(1.0, std::fabs(v[i]));
      double vp[10], vm[10], fp[6], fm[6];
      for (int k = 0; k < 10; ++k) vp[k] = vm[k] = v[k];
      vp[i] += h;
      vm[i] -= h;
      rhs(vp, s, fp);
      rhs(vm, s, fm);
      double fd = 0.0, mag = 0.0;
      for (int k = 0; k < 6; ++k) {
        fd += lam[k] * (fp[k] - fm[k]) / (2.0 * h);
        mag += std::fabs(lam[k] * (fp[k] + fm[k]) * 0.5);
      }
      double err = std::fabs(fd - got[i]) / std::fmax(1.0, std::fmax(std::fabs(fd), mag));
      if (err > worst) worst = err;
      if (err > 1e-7) {
        std::printf("case %d input %d: fd %.12g vjp %.12g (x=%g,%g,%g a=%g)\n", t, i, fd, got[i],
                    v[0], v[1], v[2], v[7]);
      }
    }
    ++cases;
  }
  std::printf("rhs_vjp cases %d worst %.3e\n", cases, worst);
  // rhs_only (the regrouped derivative of the intermediate stages) against
  // geometry + rhs_from_geometry
  double worst_only = 0.0;
  for (int t = 0; t < 2000; ++t) {
    const double s = (t & 1) ? 1.0 : -1.0;
    const double scale = (t % 4 < 2) ? 12.0 : 0.6;
    double v[10] = {scale * U(rng), scale * U(rng), scale * U(rng), U(rng), U(rng), U(rng),
                    1.0 + 0.2 * U(rng), 0.95 * U(rng), 0.15 * (U(rng) + 1.0), 1.0};
    if (v[0] * v[0] + v[1] * v[1] + v[2] * v[2] < 0.05) continue;
    Geom<double> g = geometry<double>(v[0], v[1], v[2], 1.0, v[7], v[7] * v[7], v[8], s);
    Deriv<double> a = rhs_from_geometry<double>(g, v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7], s);
    Deriv<double> b;
    rhs_only<double>(v[0], v[1], v[2], v[3], v[4], v[5], v[6], 1.0, v[7], v[7] * v[7], v[8], s, b);
    const double da[7] = {a.dt, a.dx, a.dy, a.dz, a.dpx, a.dpy, a.dpz};
    const double db[7] = {b.dt, b.dx, b.dy, b.dz, b.dpx, b.dpy, b.dpz};
    double mag = 1.0;
    for (double q : da) mag = std::fmax(mag, std::fabs(q));
    for (int i = 0; i < 7; ++i) worst_only = std::fmax(worst_only, std::fabs(da[i] - db[i]) / mag);
  }
  std::printf("rhs_only worst %.3e\n", worst_only);
  if (worst_only > 1e-12) return 1;
  const double worst_traj = trajectory_check();
  std::printf("trajectory worst %.3e\n", worst_traj);
  return (worst <= 1e-7 && worst_traj <= 1e-6) ? 0 : 1;
}
