// Reverse-mode pieces of the discrete adjoint of the trace.
//
// rhs_vjp: vector-Jacobian product of the reduced Kerr-Schild right-hand side
// (x' = p - f q l, p' = f q w + q^2/2 grad f; ks_math.h) with respect to the
// state (x, p), the per-ray energy E and the spacetime parameters (a, a^2,
// Q^2). It recomputes the forward intermediates of geometry() and
// rhs_from_geometry() and then sweeps them in reverse, line by line; the
// coordinate-time rate t' takes no part (no observable of the adjoint depends
// on t).
//
// Dual: a forward-mode number with N tangent directions, used once per ray
// for the initial-state and observable sensitivities.
#pragma once

#include <cmath>

#include "ks_math.h"
#include "trace_device.h"

namespace gl {

// Adjoint seeds and results of one RHS VJP.
template <class T>
struct RhsBar {
  T x, y, z, px, py, pz;  // d<lambda, F>/d state
  T E;                    // .../dE
  T spin, spin2, charge2; // .../d a (explicit), d a^2, d Q^2
};

// lam: adjoint of (x', y', z', px', py', pz'). Accumulates into *out.
template <class T>
GL_HD void rhs_vjp(T x, T y, T z, T px, T py, T pz, T E, T mass, T spin, T spin2, T charge2,
                   T s, const T lam[6], RhsBar<T>* out) {
  // ---- forward (geometry) ----
  const T rho2 = x * x + y * y + z * z;
  const T b = rho2 - spin2;
  const T az2 = spin2 * (z * z);
  const T disc = b * b + T(4) * az2;
  const T sinv = gl_rsqrt<T>(disc);
  const T S = disc * sinv;
  const bool pos = b >= T(0);
  const T u = S - b;  // used when b < 0
  const T r2 = pos ? T(0.5) * (b + S) : (T(2) * az2) * gl_rcp<T>(u);
  const T rinv = gl_rsqrt<T>(r2);
  const T r = r2 * rinv;
  const T D = r2 + spin2;
  const T kinv = gl_rcp<T>(D);
  const T num = T(2) * mass * r - charge2;
  const T f = num * sinv;
  const T rs = r * sinv;
  const T grx = x * rs, gry = y * rs;
  const T G = D * rinv * sinv;
  const T grz = z * G;
  const T two_sinv = T(2) * sinv;
  const T gsx = b * x * two_sinv, gsy = b * y * two_sinv;
  const T H = rho2 + spin2;
  const T gsz = z * H * two_sinv;
  const T tm = T(2) * mass;
  const T hx = tm * grx - f * gsx, hy = tm * gry - f * gsy, hz = tm * grz - f * gsz;
  const T gfx = hx * sinv, gfy = hy * sinv, gfz = hz * sinv;
  const T sr = s * r;
  const T Lx = sr * x + spin * y, Ly = sr * y - spin * x;
  const T lx = Lx * kinv, ly = Ly * kinv, lz = s * z * rinv;
  // ---- forward (rhs) ----
  const T lp = lx * px + ly * py + lz * pz;
  const T q = E + lp;
  const T fq = f * q;
  const T Ex = sr * px - spin * py, Ey = spin * px + sr * py;
  const T ex = Ex * kinv, ey = Ey * kinv, ez = s * pz * rinv;
  const T lpxy = lx * px + ly * py;
  const T A = s * (x * px + y * py) - T(2) * r * lpxy;
  const T ri2 = rinv * rinv;
  const T c = A * kinv - s * z * pz * ri2;
  const T wx = ex + c * grx, wy = ey + c * gry, wz = ez + c * grz;
  const T hq2 = T(0.5) * q * q;

  // ---- reverse (rhs) ----
  const T Dx = lam[0], Dy = lam[1], Dz = lam[2], Px = lam[3], Py = lam[4], Pz = lam[5];
  T bx = T(0), by = T(0), bz = T(0), bpx = T(0), bpy = T(0), bpz = T(0);
  T bE = T(0), ba = T(0), ba2 = T(0), bq2 = T(0);
  // dp = fq w + hq2 grad f
  T bfq = Px * wx + Py * wy + Pz * wz;
  const T bwx = Px * fq, bwy = Py * fq, bwz = Pz * fq;
  const T bhq2 = Px * gfx + Py * gfy + Pz * gfz;
  T bgfx = Px * hq2, bgfy = Py * hq2, bgfz = Pz * hq2;
  T bq = bhq2 * q;
  // w = e + c grad r
  const T bex = bwx, bey = bwy, bez = bwz;
  const T bc = bwx * grx + bwy * gry + bwz * grz;
  T bgrx = bwx * c, bgry = bwy * c, bgrz = bwz * c;
  // c = A kinv - s z pz rinv^2
  const T bA = bc * kinv;
  T bkinv = bc * A;
  bz -= bc * s * pz * ri2;
  bpz -= bc * s * z * ri2;
  T brinv = -bc * s * z * pz * T(2) * rinv;
  // A = s (x px + y py) - 2 r (lx px + ly py)
  bx += bA * s * px;
  by += bA * s * py;
  bpx += bA * (s * x - T(2) * r * lx);
  bpy += bA * (s * y - T(2) * r * ly);
  T br = -bA * T(2) * lpxy;
  T blx = -bA * T(2) * r * px;
  T bly = -bA * T(2) * r * py;
  T blz = T(0);
  // ez = s pz rinv
  bpz += bez * s * rinv;
  brinv += bez * s * pz;
  // ey = (a px + sr py) kinv
  ba += bey * px * kinv;
  bpx += bey * spin * kinv;
  T bsr = bey * py * kinv;
  bpy += bey * sr * kinv;
  bkinv += bey * Ey;
  // ex = (sr px - a py) kinv
  bsr += bex * px * kinv;
  bpx += bex * sr * kinv;
  ba -= bex * py * kinv;
  bpy -= bex * spin * kinv;
  bkinv += bex * Ex;
  // dx = p - fq l
  bpx += Dx;
  bpy += Dy;
  bpz += Dz;
  bfq -= Dx * lx + Dy * ly + Dz * lz;
  blx -= Dx * fq;
  bly -= Dy * fq;
  blz -= Dz * fq;
  // fq = f q
  T bf = bfq * q;
  bq += bfq * f;
  // q = E + lp
  bE += bq;
  const T blp = bq;
  blx += blp * px;
  bly += blp * py;
  blz += blp * pz;
  bpx += blp * lx;
  bpy += blp * ly;
  bpz += blp * lz;

  // ---- reverse (geometry) ----
  // lz = s z rinv
  bz += blz * s * rinv;
  brinv += blz * s * z;
  // ly = (sr y - a x) kinv
  bsr += bly * y * kinv;
  by += bly * sr * kinv;
  ba -= bly * x * kinv;
  bx -= bly * spin * kinv;
  bkinv += bly * Ly;
  // lx = (sr x + a y) kinv
  bsr += blx * x * kinv;
  bx += blx * sr * kinv;
  ba += blx * y * kinv;
  by += blx * spin * kinv;
  bkinv += blx * Lx;
  // sr = s r
  br += bsr * s;
  // grad f = h sinv, h = tm gr - f gs
  T bsinv = bgfx * hx + bgfy * hy + bgfz * hz;
  const T bhx = bgfx * sinv, bhy = bgfy * sinv, bhz = bgfz * sinv;
  bgrx += bhx * tm;
  bgry += bhy * tm;
  bgrz += bhz * tm;
  bf -= bhx * gsx + bhy * gsy + bhz * gsz;
  const T bgsx = -bhx * f, bgsy = -bhy * f, bgsz = -bhz * f;
  // gs = (b x, b y, z H) two_sinv
  T bb = (bgsx * x + bgsy * y) * two_sinv;
  bx += bgsx * b * two_sinv;
  by += bgsy * b * two_sinv;
  bz += bgsz * H * two_sinv;
  const T bH = bgsz * z * two_sinv;
  T brho2 = bH;
  ba2 += bH;
  bsinv += T(2) * (bgsx * b * x + bgsy * b * y + bgsz * z * H);
  // grz = z G, G = D rinv sinv
  bz += bgrz * G;
  const T bG = bgrz * z;
  T bD = bG * rinv * sinv;
  brinv += bG * D * sinv;
  bsinv += bG * D * rinv;
  // grx = x rs, gry = y rs
  bx += bgrx * rs;
  by += bgry * rs;
  const T brs = bgrx * x + bgry * y;
  // rs = r sinv
  br += brs * sinv;
  bsinv += brs * r;
  // f = num sinv, num = 2 M r - Q^2
  const T bnum = bf * sinv;
  bsinv += bf * num;
  br += bnum * tm;
  bq2 -= bnum;
  // kinv = 1 / D, D = r2 + a^2
  bD -= bkinv * kinv * kinv;
  T br2 = bD;
  ba2 += bD;
  // r = r2 rinv, rinv = r2^-1/2
  br2 += br * rinv;
  brinv += br * r2;
  br2 -= T(0.5) * brinv * rinv * ri2;
  // r2 = (b + S)/2  or  2 az2 / (S - b)
  T bS = T(0), baz2 = T(0);
  if (pos) {
    bb += T(0.5) * br2;
    bS += T(0.5) * br2;
  } else {
    const T uinv = gl_rcp<T>(u);
    baz2 += T(2) * br2 * uinv;
    const T bu = -br2 * r2 * uinv;
    bS += bu;
    bb -= bu;
  }
  // S = sqrt(disc), sinv = disc^-1/2
  T bdisc = bS * T(0.5) * sinv - T(0.5) * bsinv * sinv * sinv * sinv;
  // disc = b^2 + 4 az2
  bb += T(2) * b * bdisc;
  baz2 += T(4) * bdisc;
  // az2 = a^2 z^2
  ba2 += baz2 * z * z;
  bz += baz2 * T(2) * spin2 * z;
  // b = rho2 - a^2
  brho2 += bb;
  ba2 -= bb;
  // rho2
  bx += T(2) * x * brho2;
  by += T(2) * y * brho2;
  bz += T(2) * z * brho2;

  out->x += bx;
  out->y += by;
  out->z += bz;
  out->px += bpx;
  out->py += bpy;
  out->pz += bpz;
  out->E += bE;
  out->spin += ba;
  out->spin2 += ba2;
  out->charge2 += bq2;
}

// ---------------------------------------------------------------------------
// Frozen-step DP5 pieces of the discrete adjoint (templates in T so the host
// can check them in double against finite differences of the same map).

// DP5 coefficients a_ij (rows 2..6) and, as row 6, the fifth-order weights.
template <class T>
GL_HD T dp5_a(int i, int j) {
  const T A[6][6] = {
      {T(GL_A21), T(0), T(0), T(0), T(0), T(0)},
      {T(GL_A31), T(GL_A32), T(0), T(0), T(0), T(0)},
      {T(GL_A41), T(GL_A42), T(GL_A43), T(0), T(0), T(0)},
      {T(GL_A51), T(GL_A52), T(GL_A53), T(GL_A54), T(0), T(0)},
      {T(GL_A61), T(GL_A62), T(GL_A63), T(GL_A64), T(GL_A65), T(0)},
      {T(GL_B1), T(0), T(GL_B3), T(GL_B4), T(GL_B5), T(GL_B6)}};
  return A[i][j];
}

// Right-hand side (x', p') at y; the coordinate-time rate into *dt if asked.
template <class T>
GL_HD void rhs6(const SceneT<T>& sc, const T y[6], T E, T k[6], T* dt = nullptr) {
  Geom<T> g = geometry<T>(y[0], y[1], y[2], sc.mass, sc.spin, sc.spin2, sc.charge2, sc.chart);
  Deriv<T> d = rhs_from_geometry<T>(g, y[0], y[1], y[2], y[3], y[4], y[5], E, sc.spin, sc.chart);
  k[0] = d.dx;
  k[1] = d.dy;
  k[2] = d.dz;
  k[3] = d.dpx;
  k[4] = d.dpy;
  k[5] = d.dpz;
  if (dt) *dt = d.dt;
}

// Stage state Y_i = y + h sum_{j<i} a_ij K_j (i = 6: the fifth-order y_{n+1}).
template <class T>
GL_HD void stage_state(const T y[6], T h, const T K[7][6], int i, T Y[6]) {
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    T acc = T(0);
#pragma unroll
    for (int j = 0; j < 6; ++j)
      if (j < i) acc += dp5_a<T>(i - 1, j) * K[j][c];
    Y[c] = i == 0 ? y[c] : y[c] + h * acc;
  }
}

// One frozen-h DP5 step y -> y_{n+1}.
template <class T>
GL_HD void advance(const SceneT<T>& sc, T y[6], T E, T h) {
  T K[7][6];
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    T Y[6];
    stage_state<T>(y, h, K, i, Y);
    rhs6<T>(sc, Y, E, K[i]);
  }
  T y1[6];
  stage_state<T>(y, h, K, 6, y1);
#pragma unroll
  for (int c = 0; c < 6; ++c) y[c] = y1[c];
}

// Reverse sweep over stages i = top..0 given the stage adjoints kb; adds the
// state adjoint into yb and the parameter adjoints into acc.
template <class T>
GL_HD void reverse_stages(const SceneT<T>& sc, const T y[6], T E, T h, const T K[7][6],
                          T kb[7][6], int top, T yb[6], RhsBar<T>& acc) {
#pragma unroll
  for (int i = 6; i >= 0; --i) {
    if (i > top) continue;
    T Y[6];
    stage_state<T>(y, h, K, i, Y);
    RhsBar<T> rb = {};
    rhs_vjp<T>(Y[0], Y[1], Y[2], Y[3], Y[4], Y[5], E, sc.mass, sc.spin, sc.spin2, sc.charge2,
               sc.chart, kb[i], &rb);
    const T Yb[6] = {rb.x, rb.y, rb.z, rb.px, rb.py, rb.pz};
#pragma unroll
    for (int c = 0; c < 6; ++c) yb[c] += Yb[c];
#pragma unroll
    for (int j = 0; j < 6; ++j)
      if (j < i) {
        const T w = h * dp5_a<T>(i - 1, j);
#pragma unroll
        for (int c = 0; c < 6; ++c) kb[j][c] += w * Yb[c];
      }
    acc.E += rb.E;
    acc.spin += rb.spin;
    acc.spin2 += rb.spin2;
    acc.charge2 += rb.charge2;
  }
}

// Adjoint of one non-terminal step: lam holds dL/dy_{n+1} on entry, dL/dy_n
// on exit.
template <class T>
GL_HD void step_adjoint(const SceneT<T>& sc, const T y[6], T E, T h, T lam[6], RhsBar<T>& acc) {
  T K[7][6];
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    T Y[6];
    stage_state<T>(y, h, K, i, Y);
    rhs6<T>(sc, Y, E, K[i]);
  }
  T kb[7][6];
#pragma unroll
  for (int j = 0; j < 7; ++j) {
    const T w = j < 6 ? h * dp5_a<T>(5, j) : T(0);
#pragma unroll
    for (int c = 0; c < 6; ++c) kb[j][c] = w * lam[c];
  }
  T yb[6] = {lam[0], lam[1], lam[2], lam[3], lam[4], lam[5]};
  reverse_stages<T>(sc, y, E, h, K, kb, 5, yb, acc);
#pragma unroll
  for (int c = 0; c < 6; ++c) lam[c] = yb[c];
}

// The seven stages of the terminal step (K without, Kt with the time
// column) and the fraction theta* where z crosses zero on its dense output.
template <class T>
GL_HD T terminal_stages(const SceneT<T>& sc, const T y[6], T E, T h, T K[7][6], T Kt[7][7]) {
#pragma unroll
  for (int i = 0; i < 7; ++i) {
    T Y[6];
    stage_state<T>(y, h, K, i, Y);
    T dt;
    rhs6<T>(sc, Y, E, K[i], &dt);
    Kt[i][0] = dt;
#pragma unroll
    for (int c = 0; c < 6; ++c) Kt[i][c + 1] = K[i][c];
  }
  T y1[6];
  stage_state<T>(y, h, K, 6, y1);
  return T(z_root<double, T>(Kt, h, y[2], y1[2]));
}

// Reverse of the terminal map y -> y(theta*(y)) on the dense output, with
// theta* defined by z(theta*) = 0, seeded by lT = dL/dy(theta*). Returns
// false if z does not cross transversally.
template <class T>
GL_HD bool terminal_reverse(const SceneT<T>& sc, const T y[6], T E, T h, const T K[7][6], T th,
                            const T lT_in[6], T lam[6], RhsBar<T>& acc) {
  T ws[7], w[7];
  dense_slopes<T>(th, ws);
  dense_weights<T>(th, w);
  T yp[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    T s = T(0);
#pragma unroll
    for (int i = 0; i < 7; ++i) s += ws[i] * K[i][c];
    yp[c] = h * s;
  }
  if (!(gl_abs<T>(yp[2]) > T(0))) return false;
  T lT[6] = {lT_in[0], lT_in[1], lT_in[2], lT_in[3], lT_in[4], lT_in[5]};
  T proj = T(0);
#pragma unroll
  for (int c = 0; c < 6; ++c) proj += lT[c] * yp[c];
  lT[2] -= proj / yp[2];
  T kb[7][6];
#pragma unroll
  for (int i = 0; i < 7; ++i)
#pragma unroll
    for (int c = 0; c < 6; ++c) kb[i][c] = h * w[i] * lT[c];
  T yb[6] = {lT[0], lT[1], lT[2], lT[3], lT[4], lT[5]};
  reverse_stages<T>(sc, y, E, h, K, kb, 6, yb, acc);
#pragma unroll
  for (int c = 0; c < 6; ++c) lam[c] = yb[c];
  return true;
}

// ---------------------------------------------------------------------------
// Forward-mode dual number with N tangent directions.
template <int N>
struct Dual {
  float v;
  float d[N];
  GL_HD Dual() : v(0.0f) {
    for (int i = 0; i < N; ++i) d[i] = 0.0f;
  }
  GL_HD Dual(double x) : v(float(x)) {  // NOLINT: constants such as T(0.5)
    for (int i = 0; i < N; ++i) d[i] = 0.0f;
  }
  GL_HD static Dual var(float x, int k) {
    Dual r(0.0);
    r.v = x;
    r.d[k] = 1.0f;
    return r;
  }
};

template <int N>
GL_HD Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r;
  r.v = a.v + b.v;
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i];
  return r;
}
template <int N>
GL_HD Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r;
  r.v = a.v - b.v;
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i];
  return r;
}
template <int N>
GL_HD Dual<N> operator-(const Dual<N>& a) {
  Dual<N> r;
  r.v = -a.v;
  for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
  return r;
}
template <int N>
GL_HD Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r;
  r.v = a.v * b.v;
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
  return r;
}
template <int N>
GL_HD Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r;
  const float inv = 1.0f / b.v;
  r.v = a.v * inv;
  for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * inv;
  return r;
}
template <int N>
GL_HD Dual<N>& operator+=(Dual<N>& a, const Dual<N>& b) {
  a = a + b;
  return a;
}
template <int N>
GL_HD Dual<N>& operator-=(Dual<N>& a, const Dual<N>& b) {
  a = a - b;
  return a;
}
template <int N>
GL_HD Dual<N>& operator*=(Dual<N>& a, const Dual<N>& b) {
  a = a * b;
  return a;
}
template <int N>
GL_HD bool operator<(const Dual<N>& a, const Dual<N>& b) { return a.v < b.v; }
template <int N>
GL_HD bool operator>(const Dual<N>& a, const Dual<N>& b) { return a.v > b.v; }
template <int N>
GL_HD bool operator<=(const Dual<N>& a, const Dual<N>& b) { return a.v <= b.v; }
template <int N>
GL_HD bool operator>=(const Dual<N>& a, const Dual<N>& b) { return a.v >= b.v; }

// f(a) with f(a.v) = value and f'(a.v) = slope.
template <int N>
GL_HD Dual<N> dual_apply(const Dual<N>& a, float value, float slope) {
  Dual<N> r;
  r.v = value;
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * slope;
  return r;
}

using Dual5 = Dual<5>;

template <>
GL_HD Dual5 gl_sqrt<Dual5>(Dual5 v) {
  const float s = gl_sqrt<float>(v.v);
  return dual_apply(v, s, 0.5f / s);
}
template <>
GL_HD Dual5 gl_atan2<Dual5>(Dual5 y, Dual5 x) {
  Dual5 r;
  r.v = gl_atan2<float>(y.v, x.v);
  const float inv = 1.0f / (x.v * x.v + y.v * y.v);
  for (int i = 0; i < 5; ++i) r.d[i] = (x.v * y.d[i] - y.v * x.d[i]) * inv;
  return r;
}
template <>
GL_HD Dual5 gl_floor<Dual5>(Dual5 v) {
  return Dual5(double(gl_floor<float>(v.v)));
}

}  // namespace gl
