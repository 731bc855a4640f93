// Discrete adjoint of the f32 trace: vector-Jacobian products of per-pixel
// frequency ratios g of surface hits with respect to the spin a, the charge
// Q, the observer inclination and each pixel's sight-plane direction.
//
// Forward (k_vjp_forward): the plain image trace, plus a record of every
// accepted step size, the state every K accepted steps (checkpoints), the
// state before the terminal step, and E.
//
// Backward (k_vjp_backward), one ray per lane:
//   1. the terminal step: the surface observable's gradient (forward-mode
//      duals), the implicit dependence of the crossing fraction theta* on the
//      step (z(theta*) = 0), and the reverse sweep through the quartic dense
//      output and its seven stages;
//   2. segment by segment from the last checkpoint down: the segment's states
//      are recomputed with the frozen step sizes into a per-lane buffer, then
//      each step is reversed stage by stage (rhs_vjp, adjoint.h);
//   3. the initial state: its tangents with respect to (a, Q, inclination)
//      come from f64 central differences of the host tetrad construction,
//      combined on the device with forward-mode duals of the pixel's sight
//      momentum; the direction gradient uses the same duals.
// Step sizes are frozen (the derivative of the discrete map at fixed h), and
// the Kahan compensation and null re-projection of the forward pass are
// treated as identity maps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>

#include "../../include/gravlume.h"
#include "adjoint.h"
#include "gl_device.h"
#include "gl_plan.h"
#include "gl_scene.h"
#include "gl_types.h"
#include "trace_device.h"

namespace gl {

// Tangents of the observer state with respect to (a, Q, inclination).
struct VjpTangents {
  float obs[3][3];     // d(x, y, z)_obs / d param
  float ut[3];         // d u^t / d param
  float right[3][4];   // d e_R / d param
  float up[3][4];      // d e_U / d param
  float arrival[3][4]; // d e_A / d param
  float dq2_dq;        // d Q^2 / d Q
};

struct VjpBufs {
  float* h;       // [cap][rays] accepted step sizes
  float* ckpt;    // [ncp][6][rays] state before steps 0, K, 2K, ...
  float* term;    // [6][rays] state before the terminal step
  float* E;       // [rays]
  uint32_t* n;    // [rays] accepted steps incl. the terminal one; 0 = no gradient
  float* seg;     // [K][6][rays] backward segment states
  uint32_t rays, cap, K, ncp;
};

__device__ __forceinline__ bool vjp_pixel(const SceneDev& s, uint32_t& px, uint32_t& py) {
  px = (blockIdx.x % s.tiles_x) * 8u + (threadIdx.x & 7u);
  py = (blockIdx.x / s.tiles_x) * 8u + (threadIdx.x >> 3);
  return px < s.width && py < s.height;
}

__global__ void __launch_bounds__(64) k_vjp_forward(const SceneDev* __restrict__ scenes,
                                                    uint32_t scene, OutPlanes out, MapView map,
                                                    VjpBufs vb) {
  const SceneDev& s = scenes[scene];
  SceneT<float> sc = scene_f32(s);
  hold_in_vgprs(sc);
  uint32_t px, py;
  if (!vjp_pixel(s, px, py)) return;
  const uint32_t idx = py * s.width + px;
  const size_t rays = vb.rays;
  RayT<float> R;
  init_image_ray(s, sc, px, py, R);
  bool overflow = false, island = false;
  // the re-projection clock by increments, and a wavefront-uniform loop exit
  // (a lane whose ray ended idles masked), as in the trace's stepping loop
  const uint32_t period = sc.reproject_interval;
  uint32_t it = 1, phase = period > 1u ? 1u : 0u;
  for (;;) {
    const bool live = R.term == kTermActive;
    if (!__any(live)) break;
    if (!live) continue;
    // where the trace would hand the ray to its f64 finisher, this forward
    // goes on in f32: such a ray is not the trace's, and gets no gradient
    island = island || wants_defer(sc, R);
    const float pre[6] = {R.x - R.cx, R.y - R.cy, R.z - R.cz, R.px, R.py, R.pz};
    const float h = R.h;
    const uint32_t acc = R.accepted;
    step<float>(sc, R, s.omega, period != 0u && phase == 0u, drift_now(it));
    ++it;
    phase = phase + 1u == period ? 0u : phase + 1u;
    if (R.accepted != acc) {
      if (acc < vb.cap) {
        vb.h[size_t(acc) * rays + idx] = h;
        if (acc % vb.K == 0u) {
#pragma unroll
          for (int c = 0; c < 6; ++c) vb.ckpt[(size_t(acc / vb.K) * 6 + c) * rays + idx] = pre[c];
        }
      } else {
        overflow = true;
      }
      if (R.term != kTermActive) {
#pragma unroll
        for (int c = 0; c < 6; ++c) vb.term[size_t(c) * rays + idx] = pre[c];
      }
    }
  }
  vb.E[idx] = R.E;
  vb.n[idx] = (R.term == kTermSurface && !overflow && !island) ? R.accepted : 0u;
  write_ray(R, out, idx, s, map, pixel_source(px, py), no_refine(), 0u);
  if (R.term == kTermSurface && (overflow || island) && out.status)
    out.status[idx] |= GL_STATUS_NO_ADJOINT;
}

// Adjoint of the terminal (surface-crossing) step, seeded by dL/d(r, phi_s,
// g) of the surface observable. Returns false when the observable is not
// differentiable there.
__device__ __forceinline__ bool terminal_adjoint(const SceneT<float>& sc, const SceneDev& s,
                                                 const float y[6], float E, float h, float wr,
                                                 float wphi, float wg, float lam[6],
                                                 RhsBar<float>& acc) {
  float K[7][6], Kt[7][7];
  const float th = terminal_stages<float>(sc, y, E, h, K, Kt);
  float yT[6], dtl;
  dense_state<float>(Kt, h, y, th, yT, dtl);
  // observable g(r, E, L_z, a, Q^2) by forward-mode duals
  Geom<float> gT = geometry<float>(yT[0], yT[1], yT[2], sc.mass, sc.spin, sc.spin2, sc.charge2,
                                   sc.chart);
  const float lz = yT[0] * yT[4] - yT[1] * yT[3];
  Dual5 o[3];
  if (!surface_observable<Dual5>(Dual5(1.0), Dual5::var(sc.spin, 3), Dual5::var(sc.charge2, 4),
                                 Dual5(double(sc.chart)), Dual5(double(s.omega)),
                                 Dual5(double(yT[0])), Dual5(double(yT[1])), Dual5::var(gT.r, 0),
                                 Dual5::var(E, 1), Dual5::var(lz, 2), o))
    return false;
  // g depends on (r, E, L_z, a, Q^2); phi_s = atan2(y, x) - atan2(s a, r) on
  // (x, y) directly and on (r, a) through the duals; r on the position and a^2
  const float glz = o[2].d[2];
  const float cr = wr + wphi * float(o[1].d[0]) + wg * float(o[2].d[0]);  // dL/dr in total
  const float rr = gT.r;
  const float rho2 = yT[0] * yT[0] + yT[1] * yT[1] + yT[2] * yT[2];
  const float dr_da2 = -(rr * rr - yT[2] * yT[2]) /
                       (4.0f * rr * rr * rr - 2.0f * rr * (rho2 - sc.spin2));
  const float rxy2 = yT[0] * yT[0] + yT[1] * yT[1];
  const float px = -wphi * yT[1] / rxy2, py = wphi * yT[0] / rxy2;  // d(atan2(y, x))
  const float lT[6] = {cr * gT.grx + wg * glz * yT[4] + px, cr * gT.gry - wg * glz * yT[3] + py,
                       cr * gT.grz, -wg * glz * yT[1], wg * glz * yT[0], 0.0f};
  acc.E += wg * float(o[2].d[1]);
  acc.spin += wg * float(o[2].d[3]) + wphi * float(o[1].d[3]);
  acc.spin2 += cr * dr_da2;
  acc.charge2 += wg * float(o[2].d[4]) + wphi * float(o[1].d[4]);
  return terminal_reverse<float>(sc, y, E, h, K, th, lT, lam, acc);
}

struct Grad5 {
  float a, q, incl, sx, sy;
};

// The sight momentum of pixel (px, py) in forward-mode duals over
// (a, Q, inclination, s_x, s_y), contracted with the initial-state adjoint
// (lam, lamE). Out of line: the duals' registers stay out of the trace loop.
__device__ __noinline__ Grad5 initial_contract(const SceneDev& s, const VjpTangents* tg,
                                               uint32_t px, uint32_t py, float l0, float l1,
                                               float l2, float l3, float l4, float l5,
                                               float lE) {
  const float nx = 2.0f * (float(px) + s.sub_x) / float(s.width) - 1.0f;
  const float ny = 1.0f - 2.0f * (float(py) + s.sub_y) / float(s.height);
  const Dual5 sx = Dual5::var(s.aspect * s.tan_half_fov * nx, 3);
  const Dual5 sy = Dual5::var(s.tan_half_fov * ny, 4);
  auto tang = [](float v, const float t[3][4], int c) {
    Dual5 r(0.0);
    r.v = v;
    for (int k = 0; k < 3; ++k) r.d[k] = t[k][c];
    return r;
  };
  const Dual5 inv = Dual5(1.0) / gl_sqrt<Dual5>(Dual5(1.0) + sx * sx + sy * sy);
  Dual5 d[4];
  for (int i = 0; i < 4; ++i)
    d[i] = (sx * tang(s.right[i], tg->right, i) + sy * tang(s.up[i], tg->up, i) -
            tang(s.arrival[i], tg->arrival, i)) *
           inv;
  Dual5 ut(0.0);
  ut.v = s.u_t;
  for (int k = 0; k < 3; ++k) ut.d[k] = tg->ut[k];
  const Dual5 om(double(s.omega));
  const Dual5 pu[4] = {om * (ut - d[0]), -(om * d[1]), -(om * d[2]), -(om * d[3])};
  Dual5 ob[3];
  const float ov[3] = {s.obs_x, s.obs_y, s.obs_z};
  for (int c = 0; c < 3; ++c) {
    ob[c] = Dual5(0.0);
    ob[c].v = ov[c];
    for (int k = 0; k < 3; ++k) ob[c].d[k] = tg->obs[k][c];
  }
  const Dual5 spin = Dual5::var(s.spin, 0);
  Dual5 q2(0.0);
  q2.v = s.charge2;
  q2.d[1] = tg->dq2_dq;
  Geom<Dual5> g = geometry<Dual5>(ob[0], ob[1], ob[2], Dual5(double(s.mass)), spin, spin * spin,
                                  q2, Dual5(double(s.chart)));
  const Dual5 lp = pu[0] + g.lx * pu[1] + g.ly * pu[2] + g.lz * pu[3];
  const Dual5 fl = g.f * lp;
  const Dual5 E = pu[0] - fl;  // E = -p_t, p_t = -pu^t + f (l.pu)
  const Dual5 p[3] = {pu[1] + fl * g.lx, pu[2] + fl * g.ly, pu[3] + fl * g.lz};
  const float lam[6] = {l0, l1, l2, l3, l4, l5};
  float out[5];
  for (int k = 0; k < 5; ++k) {
    float v = lE * E.d[k];
    for (int c = 0; c < 3; ++c) v += lam[c] * ob[c].d[k] + lam[c + 3] * p[c].d[k];
    out[k] = v;
  }
  return Grad5{out[0], out[1], out[2], out[3], out[4]};
}

// Register budget of two waves per SIMD (256 VGPRs): its natural 283 left
// one wave per SIMD with nothing to hide the adjoint's latency; 23 spills
// cost less than the second wave gains (backward about 7% faster).
__global__ void __launch_bounds__(64, 2) k_vjp_backward(const SceneDev* __restrict__ scenes,
                                                     uint32_t scene, VjpBufs vb,
                                                     const float* __restrict__ grad_r,
                                                     const float* __restrict__ grad_phi,
                                                     const float* __restrict__ grad_g,
                                                     float* __restrict__ grad_dir,
                                                     const VjpTangents* __restrict__ tg,
                                                     double* __restrict__ grad_acc) {
  const SceneDev& s = scenes[scene];
  const SceneT<float> sc = scene_f32(s);
  uint32_t px, py;
  const bool valid = vjp_pixel(s, px, py);
  const size_t rays = vb.rays;
  const uint32_t idx = valid ? py * s.width + px : 0u;
  const uint32_t N = valid ? vb.n[idx] : 0u;
  const float wr = valid && grad_r ? grad_r[idx] : 0.0f;
  const float wphi = valid && grad_phi ? grad_phi[idx] : 0.0f;
  const float wg = valid && grad_g ? grad_g[idx] : 0.0f;
  float gp[3] = {0.0f, 0.0f, 0.0f};
  float gd[2] = {0.0f, 0.0f};
  if (N > 0u && (wr != 0.0f || wphi != 0.0f || wg != 0.0f)) {
    const float E = vb.E[idx];
    RhsBar<float> acc = {};
    float y[6], lam[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) y[c] = vb.term[size_t(c) * rays + idx];
    if (terminal_adjoint(sc, s, y, E, vb.h[size_t(N - 1) * rays + idx], wr, wphi, wg, lam, acc)) {
      if (N >= 2u) {
        for (int m = int((N - 2u) / vb.K); m >= 0; --m) {
          const uint32_t n0 = uint32_t(m) * vb.K;
          const uint32_t n1 = min(n0 + vb.K, N - 1u);
          float yy[6];
#pragma unroll
          for (int c = 0; c < 6; ++c) yy[c] = vb.ckpt[(size_t(m) * 6 + c) * rays + idx];
          for (uint32_t n = n0; n < n1; ++n) {
#pragma unroll
            for (int c = 0; c < 6; ++c) vb.seg[(size_t(n - n0) * 6 + c) * rays + idx] = yy[c];
            advance<float>(sc, yy, E, vb.h[size_t(n) * rays + idx]);
          }
          for (uint32_t n = n1; n-- > n0;) {
#pragma unroll
            for (int c = 0; c < 6; ++c) yy[c] = vb.seg[(size_t(n - n0) * 6 + c) * rays + idx];
            step_adjoint<float>(sc, yy, E, vb.h[size_t(n) * rays + idx], lam, acc);
          }
        }
      }
      const Grad5 g5 = initial_contract(s, tg, px, py, lam[0], lam[1], lam[2], lam[3], lam[4],
                                        lam[5], acc.E);
      gp[0] = g5.a + acc.spin + 2.0f * sc.spin * acc.spin2;
      gp[1] = g5.q + acc.charge2 * tg->dq2_dq;
      gp[2] = g5.incl;
      gd[0] = g5.sx;
      gd[1] = g5.sy;
    }
  }
  if (valid && grad_dir) {
    grad_dir[idx] = gd[0];
    grad_dir[rays + idx] = gd[1];
  }
#pragma unroll
  for (int k = 0; k < 3; ++k)
    for (int off = 32; off > 0; off >>= 1) gp[k] += __shfl_xor(gp[k], off);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < 3; ++k) atomicAdd(&grad_acc[k], double(gp[k]));
  }
}

// A seed of one on every ray (the adjoint footprint's two passes).
__global__ void __launch_bounds__(256) k_fill_ones(float* __restrict__ v, uint32_t n) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i < n) v[i] = 1.0f;
}

// The source-chart footprint from the two adjoint passes: with
// dr = d r / d(s_x, s_y) and dphi = d phi_s / d(s_x, s_y) (planar, 2 x n)
// and the pixel-to-sight-plane scales dsx, dsy,
//   J = (dr/dx, dr/dy, r dphi/dx, r dphi/dy)   [M per pixel],
// the rows (radial, local arc) of surface_footprint_v1 (footprint.rs:74-151)
// without its four neighbour rays. A pixel resolves when the forward ray
// ended unambiguously on the surface (no flags, candidate mask exactly the
// surface bit) within the adjoint's step capacity; parity and meta as in
// k_surface_footprint.
__global__ void __launch_bounds__(256) k_adjoint_footprint(
    VjpBufs vb, const uint32_t* __restrict__ status, const float* __restrict__ observable,
    const float* __restrict__ dr, const float* __restrict__ dphi, float dsx, float dsy,
    float* __restrict__ jac, uint32_t* __restrict__ meta) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  const uint32_t n = vb.rays;
  if (i >= n) return;
  const uint32_t st = status[i];
  const uint32_t term = st & 15u, cand = (st >> 4) & 15u, flags = (st >> 8) & 255u;
  float J[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  uint32_t resolved = 0u, parity = 0u;
  if (vb.n[i] > 0u && term == kTermSurface && flags == 0u && cand == GL_EVENT_SURFACE) {
    const float r = observable[i];
    J[0] = dr[i] * dsx;
    J[1] = dr[n + i] * dsy;
    J[2] = r * dphi[i] * dsx;
    J[3] = r * dphi[n + i] * dsy;
    const float det = J[0] * J[3] - J[1] * J[2];
    const float norm2 = J[0] * J[0] + J[1] * J[1] + J[2] * J[2] + J[3] * J[3];
    const float floor_ = 64.0f * 1.1920929e-7f * norm2;
    parity = det > floor_ ? 1u : (det < -floor_ ? 2u : 3u);
    resolved = (norm2 == norm2 && norm2 <= 3.0e38f) ? 1u : 0u;
    if (!resolved) {
      J[0] = J[1] = J[2] = J[3] = 0.0f;
      parity = 0u;
    }
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) jac[4 * size_t(i) + c] = J[c];
  meta[i] = resolved | (parity << 1) | (term << 8) | (flags << 16);
}

}  // namespace gl

// ===========================================================================
// Host side

using namespace gl;

struct gl_vjp_state {
  gl_trace_plan* plan = nullptr;
  uint32_t scene = 0;
  uint32_t tiles = 0;
  VjpBufs vb = {};
  VjpTangents* d_tangents = nullptr;
  double* d_grad = nullptr;
  double* h_grad = nullptr;  // pinned
  // the adjoint footprint's seed and direction gradients (allocated on use)
  float* d_seed = nullptr;
  float* d_dr = nullptr;
  float* d_dphi = nullptr;
};

namespace {

void release(gl_vjp_state* st) {
  if (!st) return;
  void* dev[] = {st->vb.h,       st->vb.ckpt,  st->vb.term, st->vb.E, st->vb.n, st->vb.seg,
                 st->d_tangents, st->d_grad,   st->d_seed,  st->d_dr, st->d_dphi};
  for (void* p : dev)
    if (p) (void)hipFree(p);
  if (st->h_grad) (void)hipHostFree(st->h_grad);
  delete st;
}

// Observer state of a scene re-packed at (a, Q, inclination); the observer
// keeps its oblate radius and azimuth.
int observer_at(const gl_scene_desc& d, double r, double phi, double a, double q, double th,
                SceneHost* out, std::string* err) {
  gl_scene_desc e = d;
  const double s = d.chart >= 0 ? 1.0 : -1.0;
  e.spin = a;
  e.charge = q;
  e.observer_txyz[1] = (r * std::cos(phi) - s * a * std::sin(phi)) * std::sin(th);
  e.observer_txyz[2] = (r * std::sin(phi) + s * a * std::cos(phi)) * std::sin(th);
  e.observer_txyz[3] = r * std::cos(th);
  SceneDev sd;
  return prepare_scene(e, &sd, out, err);
}

// f64 central differences of the observer position and tetrad.
gl_status observer_tangents(const SceneHost& hs, double inclination, VjpTangents* t) {
  const gl_scene_desc& d = hs.desc;
  const double s = d.chart >= 0 ? 1.0 : -1.0;
  const double ox = d.observer_txyz[1], oy = d.observer_txyz[2], oz = d.observer_txyz[3];
  const double r = geometry<double>(ox, oy, oz, d.mass, d.spin, d.spin * d.spin,
                                    d.charge * d.charge, s).r;
  const double th = std::acos(std::fmax(-1.0, std::fmin(1.0, oz / r)));
  const double phi = std::atan2(oy, ox) - std::atan2(s * d.spin, r);
  if (!(std::fabs(th - inclination) <= 1e-6))
    return fail(GL_ERR_INVALID_ARGUMENT,
                "inclination does not match the scene's observer position (polar angle " +
                    std::to_string(th) + ")");
  const double step = 1e-6;
  for (int k = 0; k < 3; ++k) {
    SceneHost hp, hm;
    std::string err;
    double ap = d.spin, am = d.spin, qp = d.charge, qm = d.charge, tp = th, tm = th;
    if (k == 0) ap += step, am -= step;
    if (k == 1) qp += step, qm -= step;
    if (k == 2) tp += step, tm -= step;
    if (observer_at(d, r, phi, ap, qp, tp, &hp, &err) || observer_at(d, r, phi, am, qm, tm, &hm, &err))
      return fail(GL_ERR_DOMAIN, "perturbed scene for the observer tangents: " + err);
    // The right leg is the one direction left by (u, up, arrival), whichever
    // axis seeds it; the image-up fallback is not: a different fallback axis
    // on the two sides of the difference would make it meaningless
    if (hp.up_axis != hs.up_axis || hm.up_axis != hs.up_axis)
      return fail(GL_ERR_DOMAIN,
                  "the observer's image-up axis changes within the tangent step; choose an up "
                  "hint that is not along the line of sight");
    const double inv = 1.0 / (2.0 * step);
    for (int c = 0; c < 3; ++c)
      t->obs[k][c] = float((hp.desc.observer_txyz[c + 1] - hm.desc.observer_txyz[c + 1]) * inv);
    t->ut[k] = float((hp.u[0] - hm.u[0]) * inv);
    for (int c = 0; c < 4; ++c) {
      t->right[k][c] = float((hp.right[c] - hm.right[c]) * inv);
      t->up[k][c] = float((hp.up[c] - hm.up[c]) * inv);
      t->arrival[k][c] = float((hp.arrival[c] - hm.arrival[c]) * inv);
    }
  }
  t->dq2_dq = float(2.0 * d.charge);
  return GL_OK;
}

}  // namespace

extern "C" {

gl_status gl_vjp_forward(gl_trace_plan* p, uint32_t scene, double inclination,
                         uint32_t checkpoint_interval, const gl_ray_outputs* out, void* stream,
                         gl_vjp_state** out_state) {
  if (out_state) *out_state = nullptr;
  if (!p || !out || !out_state) return fail(GL_ERR_INVALID_ARGUMENT, "plan, outputs and out_state are required");
  if (scene >= p->scenes.size()) return fail(GL_ERR_INVALID_ARGUMENT, "scene index out of range");
  if (checkpoint_interval == 0 || checkpoint_interval > 1024)
    return fail(GL_ERR_INVALID_ARGUMENT, "checkpoint_interval must lie in [1, 1024]");
  const SceneDev& sd = p->scenes[scene];
  const SceneHost& hs = p->host[scene];
  if (!hs.desc.has_surface) return fail(GL_ERR_INVALID_ARGUMENT, "the adjoint needs a scene with a surface");
  const size_t rays = size_t(sd.width) * sd.height;
  if (out->count < rays) return fail(GL_ERR_INVALID_ARGUMENT, "output planes hold fewer rays than the scene");
  VjpTangents tan;
  gl_status s = observer_tangents(hs, inclination, &tan);
  if (s) return s;
  gl_vjp_state* st = new gl_vjp_state();
  st->plan = p;
  st->scene = scene;
  st->tiles = sd.tiles_x * ((sd.height + 7u) / 8u);
  VjpBufs& vb = st->vb;
  vb.rays = uint32_t(rays);
  vb.K = checkpoint_interval;
  vb.cap = std::min<uint32_t>(sd.max_accepted, 1024u);
  vb.ncp = (vb.cap + vb.K - 1u) / vb.K;
  const size_t seg_steps = std::min<uint32_t>(vb.K, vb.cap);
  hipError_t e = hipSuccess;
  auto alloc = [&](void** ptr, size_t bytes) {
    if (e == hipSuccess) e = hipMalloc(ptr, bytes);
  };
  alloc((void**)&vb.h, size_t(vb.cap) * rays * sizeof(float));
  alloc((void**)&vb.ckpt, size_t(vb.ncp) * 6 * rays * sizeof(float));
  alloc((void**)&vb.term, 6 * rays * sizeof(float));
  alloc((void**)&vb.E, rays * sizeof(float));
  alloc((void**)&vb.n, rays * sizeof(uint32_t));
  alloc((void**)&vb.seg, seg_steps * 6 * rays * sizeof(float));
  alloc((void**)&st->d_tangents, sizeof(VjpTangents));
  alloc((void**)&st->d_grad, 3 * sizeof(double));
  if (e == hipSuccess) e = hipHostMalloc((void**)&st->h_grad, 3 * sizeof(double), hipHostMallocDefault);
  if (e == hipSuccess) e = hipMemcpy(st->d_tangents, &tan, sizeof(VjpTangents), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    release(st);
    return fail(e == hipErrorOutOfMemory ? GL_ERR_CAPACITY : GL_ERR_DEVICE,
                std::string("adjoint buffers: ") + hipGetErrorString(e));
  }
  hipStream_t hs_ = (hipStream_t)stream;
  MapView map;
  map.rgb = p->d_map;
  map.w = p->map_w;
  map.h = p->map_h;
  const OutPlanes planes = to_planes(*out);
  s = join_side(p, hs_);
  if (s) {
    release(st);
    return s;
  }
  plan_mark(p, 0, hs_);
  hipLaunchKernelGGL(k_vjp_forward, dim3(st->tiles), dim3(64), 0, hs_, p->d_scenes, scene, planes,
                     map, vb);
  plan_mark(p, 1, hs_);
  e = hipGetLastError();
  if (e != hipSuccess) {
    release(st);
    return fail(GL_ERR_DEVICE, std::string("k_vjp_forward: ") + hipGetErrorString(e));
  }
  *out_state = st;
  return GL_OK;
}

gl_status gl_vjp_backward(gl_vjp_state* st, const float* grad_g, double* grad_params,
                          float* grad_direction, void* stream) {
  if (!grad_g) return fail(GL_ERR_INVALID_ARGUMENT, "grad_g is required");
  return gl_vjp_backward_observables(st, nullptr, nullptr, grad_g, grad_params, grad_direction,
                                     stream);
}

gl_status gl_vjp_backward_observables(gl_vjp_state* st, const float* grad_r,
                                      const float* grad_phi, const float* grad_g,
                                      double* grad_params, float* grad_direction, void* stream) {
  if (!st || !grad_params)
    return fail(GL_ERR_INVALID_ARGUMENT, "state and grad_params are required");
  if (!grad_r && !grad_phi && !grad_g)
    return fail(GL_ERR_INVALID_ARGUMENT, "at least one observable seed is required");
  hipStream_t hs_ = (hipStream_t)stream;
  GL_HIP(hipMemsetAsync(st->d_grad, 0, 3 * sizeof(double), hs_));
  plan_mark(st->plan, 0, hs_);
  hipLaunchKernelGGL(k_vjp_backward, dim3(st->tiles), dim3(64), 0, hs_, st->plan->d_scenes,
                     st->scene, st->vb, grad_r, grad_phi, grad_g, grad_direction, st->d_tangents,
                     st->d_grad);
  plan_mark(st->plan, 1, hs_);
  GL_HIP(hipGetLastError());
  GL_HIP(hipMemcpyAsync(st->h_grad, st->d_grad, 3 * sizeof(double), hipMemcpyDeviceToHost, hs_));
  GL_HIP(hipStreamSynchronize(hs_));
  for (int k = 0; k < 3; ++k) grad_params[k] = st->h_grad[k];
  return GL_OK;
}

gl_status gl_vjp_surface_footprint(gl_vjp_state* st, const gl_ray_outputs* out, float* jacobian,
                                   uint32_t* meta, void* stream) {
  if (!st || !out || !jacobian || !meta)
    return fail(GL_ERR_INVALID_ARGUMENT, "state, outputs, jacobian and meta are required");
  if (!out->status || !out->observable)
    return fail(GL_ERR_INVALID_ARGUMENT, "the forward's status and observable planes are required");
  const uint32_t n = st->vb.rays;
  if (out->count < n) return fail(GL_ERR_INVALID_ARGUMENT, "output planes hold fewer rays than the scene");
  hipStream_t hs_ = (hipStream_t)stream;
  if (!st->d_seed) {
    hipError_t e = hipMalloc((void**)&st->d_seed, size_t(n) * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&st->d_dr, 2 * size_t(n) * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&st->d_dphi, 2 * size_t(n) * sizeof(float));
    if (e != hipSuccess) {
      (void)hipFree(st->d_seed);
      (void)hipFree(st->d_dr);
      (void)hipFree(st->d_dphi);
      st->d_seed = st->d_dr = st->d_dphi = nullptr;
      return fail(e == hipErrorOutOfMemory ? GL_ERR_CAPACITY : GL_ERR_DEVICE,
                  std::string("adjoint footprint buffers: ") + hipGetErrorString(e));
    }
  }
  const dim3 fill((n + 255u) / 256u);
  hipLaunchKernelGGL(k_fill_ones, fill, dim3(256), 0, hs_, st->d_seed, n);
  GL_HIP(hipGetLastError());
  // seed r, then phi_s; the parameter gradients they accumulate are unused
  hipLaunchKernelGGL(k_vjp_backward, dim3(st->tiles), dim3(64), 0, hs_, st->plan->d_scenes,
                     st->scene, st->vb, st->d_seed, (const float*)nullptr, (const float*)nullptr,
                     st->d_dr, st->d_tangents, st->d_grad);
  GL_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_vjp_backward, dim3(st->tiles), dim3(64), 0, hs_, st->plan->d_scenes,
                     st->scene, st->vb, (const float*)nullptr, st->d_seed, (const float*)nullptr,
                     st->d_dphi, st->d_tangents, st->d_grad);
  GL_HIP(hipGetLastError());
  const SceneDev& sd = st->plan->scenes[st->scene];
  const float dsx = 2.0f * sd.aspect * sd.tan_half_fov / float(sd.width);  // d s_x / d pixel x
  const float dsy = -2.0f * sd.tan_half_fov / float(sd.height);            // d s_y / d pixel y
  hipLaunchKernelGGL(k_adjoint_footprint, fill, dim3(256), 0, hs_, st->vb, out->status,
                     out->observable, st->d_dr, st->d_dphi, dsx, dsy, jacobian, meta);
  GL_HIP(hipGetLastError());
  return GL_OK;
}

gl_status gl_vjp_destroy(gl_vjp_state* st) {
  release(st);
  return GL_OK;
}

}  // extern "C"
