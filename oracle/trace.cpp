// Oracle (test infrastructure only): f64 Dormand-Prince 5(4) null-geodesic
// trace with dense event localization, branch key, Kahan travel time and
// invariant drifts. Restates crates/gravlume-reference/src/{integrator,
// policy,events,geodesic,surface}.rs.
#include <algorithm>
#include <cmath>
#include <limits>

#include "oracle.hpp"

namespace orc {

int singularity_guard_residual(const Spacetime& st, const double* xyz, double guard, double* out);

namespace {

constexpr int kStages = 7;

// Dormand-Prince 5(4) tableau (integrator.rs:6-98; Dormand & Prince 1980).
const double A[kStages][kStages] = {
    {0, 0, 0, 0, 0, 0, 0},
    {1.0 / 5.0, 0, 0, 0, 0, 0, 0},
    {3.0 / 40.0, 9.0 / 40.0, 0, 0, 0, 0, 0},
    {44.0 / 45.0, -56.0 / 15.0, 32.0 / 9.0, 0, 0, 0, 0},
    {19372.0 / 6561.0, -25360.0 / 2187.0, 64448.0 / 6561.0, -212.0 / 729.0, 0, 0, 0},
    {9017.0 / 3168.0, -355.0 / 33.0, 46732.0 / 5247.0, 49.0 / 176.0, -5103.0 / 18656.0, 0, 0},
    {35.0 / 384.0, 0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0, 0},
};
const double B5[kStages] = {35.0 / 384.0, 0, 500.0 / 1113.0, 125.0 / 192.0,
                            -2187.0 / 6784.0, 11.0 / 84.0, 0};
const double B4[kStages] = {5179.0 / 57600.0, 0, 7571.0 / 16695.0, 393.0 / 640.0,
                            -92097.0 / 339200.0, 187.0 / 2100.0, 1.0 / 40.0};
// Quartic continuous extension (Shampine 1986 coefficients as used by
// integrator.rs:60-98): b_i(theta) = sum_k D[i][k] theta^(k+1).
const double D[kStages][4] = {
    {1.0, -8048581381.0 / 2820520608.0, 8663915743.0 / 2820520608.0,
     -12715105075.0 / 11282082432.0},
    {0, 0, 0, 0},
    {0, 131558114200.0 / 32700410799.0, -68118460800.0 / 10900136933.0,
     87487479700.0 / 32700410799.0},
    {0, -1754552775.0 / 470086768.0, 14199869525.0 / 1410260304.0,
     -10690763975.0 / 1880347072.0},
    {0, 127303824393.0 / 49829197408.0, -318862633887.0 / 49829197408.0,
     701980252875.0 / 199316789632.0},
    {0, -282668133.0 / 205662961.0, 2019193451.0 / 616988883.0, -1453857185.0 / 822651844.0},
    {0, 40617522.0 / 29380423.0, -110615467.0 / 29380423.0, 69997945.0 / 29380423.0},
};

struct Dense {
  S8 start;
  double step;
  S8 k[kStages];
  double derivative(int comp, const double pw[4]) const {
    double s = 0.0;
    for (int i = 0; i < kStages; ++i) {
      double c = 0.0;
      for (int j = 0; j < 4; ++j) c += D[i][j] * pw[j];
      s += c * k[i][comp];
    }
    return s;
  }
  // DenseOutput::evaluate (integrator.rs:129-137).
  S8 evaluate(double th) const {
    double t2 = th * th;
    double pw[4] = {th, t2, t2 * th, t2 * t2};
    S8 r;
    for (int c = 0; c < 8; ++c) r[c] = std::fma(step, derivative(c, pw), start[c]);
    return r;
  }
  double time_increment(double th) const {
    double t2 = th * th;
    double pw[4] = {th, t2, t2 * th, t2 * t2};
    return step * derivative(0, pw);
  }
};

double weighted(const S8* k, const double* w, int comp) {
  double s = 0.0;
  for (int i = 0; i < kStages; ++i) s += w[i] * k[i][comp];
  return s;
}

struct Attempt {
  S8 end, error, end_derivative;
  Dense dense;
};

// attempt_step (integrator.rs:153-190). Returns 0 or the GeometryError of the
// failing stage; *evals receives the RHS evaluations performed.
int attempt_step(const Spacetime& st, const S8& start, double h, const S8& d0, Attempt* at,
                 int* evals) {
  S8* k = at->dense.k;
  k[0] = d0;
  for (int si = 1; si < kStages; ++si) {
    S8 ys;
    for (int c = 0; c < 8; ++c) {
      double w = 0.0;
      for (int j = 0; j < si; ++j) w += A[si][j] * k[j][c];
      ys[c] = std::fma(h, w, start[c]);
    }
    int err = hamiltonian_rhs(st, ys, &k[si]);
    if (err) {
      *evals = si;  // the failing evaluation counts (integrator.rs:170-173)
      return err;
    }
  }
  *evals = 6;
  S8 y4;
  for (int c = 0; c < 8; ++c) {
    at->end[c] = std::fma(h, weighted(k, B5, c), start[c]);
    y4[c] = std::fma(h, weighted(k, B4, c), start[c]);
    at->error[c] = at->end[c] - y4[c];
  }
  at->error[0] = h * weighted(k, B5, 0) - h * weighted(k, B4, 0);
  at->end_derivative = k[6];
  at->dense.start = start;
  at->dense.step = h;
  return 0;
}

// ReferencePolicy::error_norm (policy.rs:137-158).
double error_norm(const Policy& p, const S8& s, const S8& e, const S8& err) {
  double m = 0.0;
  for (int i = 0; i < 8; ++i) {
    double rt = i < 4 ? p.pos_rtol : p.mom_rtol;
    double at = i < 4 ? p.pos_atol : p.mom_atol;
    double v = std::fabs(err[i]) / std::fma(rt, std::fmax(std::fabs(s[i]), std::fabs(e[i])), at);
    m = std::fmax(m, v);
  }
  return m;
}

// ReferencePolicy::next_step_magnitude (policy.rs:160-173); `cap` is the
// policy's maximum step (the reference's), or less under the optional
// relative cap (step_cap).
double next_step(const Policy& p, double cur, double en, bool accepted, double cap) {
  double maxf = accepted ? p.max_factor : 1.0;
  double f = en == 0.0 ? maxf : std::clamp(p.safety * std::pow(en, -0.2), p.min_factor, maxf);
  return std::clamp(cur * f, p.minimum_step, cap);
}

// The maximum step from state y: the policy's, or with a positive
// maximum_step_per_radius also that multiple of y's oblate radius (never
// below the minimum step).
double step_cap(const Policy& p, const Spacetime& st, const S8& y) {
  double cap = p.maximum_step;
  if (p.maximum_step_per_radius > 0.0) {
    Geometry g;
    if (geometry(st, &y[1], &g) == 0 && std::isfinite(g.radius))
      cap = std::fmin(cap, p.maximum_step_per_radius * g.radius);
  }
  return std::fmax(cap, p.minimum_step);
}

enum Kind { kSing = 0, kHor = 1, kSurf = 2, kEsc = 3 };

struct CompensatedSum {  // geodesic.rs:508-528
  double sum = 0.0, corr = 0.0;
  void add(double v) {
    double next = sum + v;
    corr += std::fabs(sum) >= std::fabs(v) ? (sum - next) + v : (v - next) + sum;
    sum = next;
  }
  double total() const { return sum + corr; }
};

bool same_sign(double a, double b) { return std::signbit(a) == std::signbit(b); }

// crosses (geodesic.rs:702-708).
bool crosses(int kind, double s, double e) {
  switch (kind) {
    case kSing:
    case kHor:
      return s > 0.0 && e <= 0.0;
    case kSurf:
      return (s < 0.0 && e >= 0.0) || (s > 0.0 && e <= 0.0);
    default:
      return s < 0.0 && e >= 0.0;
  }
}

// clears_arming_band (geodesic.rs:567-573, events.rs:5-7).
bool clears_band(int kind, double v, double band) {
  switch (kind) {
    case kSing:
    case kHor:
      return v > band;
    case kSurf:
      return std::fabs(v) > band;
    default:
      return v < -band;
  }
}

double wrapped_angle_difference(double cur, double prev) {  // geodesic.rs:714-716
  double x = cur - prev + M_PI;
  double m = 2.0 * M_PI;
  double r = std::fmod(x, m);
  if (r < 0.0) r += m;
  return r - M_PI;
}

double relative_drift(double init, double cur) {  // geodesic.rs:718-720
  return std::fabs(cur - init) / std::fmax(std::fabs(init), 1.0);
}

int failure_of(int geom_err) {
  switch (geom_err) {
    case kGeomNonFinite:
      return kFailNonFinite;
    case kGeomRingSingularity:
      return kFailRingSingularity;
    case kGeomChartBoundary:
      return kFailChartBoundary;
    default:
      return kFailInvalidDenominator;
  }
}

struct Root {
  int kind;
  double theta;
  S8 state;
  double bracket;
  double residual;
};

// TraceExecution (geodesic.rs:71-506).
class Execution {
 public:
  Execution(const Spacetime& st, const Policy& pol, const Events& ev, const S8& y0, double dir)
      : st_(st), pol_(pol), ev_(ev), state_(y0), sign_(dir) {
    step_ = sign_ * pol_.initial_step;
    Invariants inv;
    has_initial_ = invariants(st_, y0, &inv) == 0;
    initial_ = inv;
    for (int k = 0; k < 4; ++k) {
      double v;
      armed_[k] = value(k, y0, &v) == 0 && clears_band(k, v, pol_.arming_band);
    }
    polar_side_ = y0[3] > 0.0 ? 2 : (y0[3] < 0.0 ? 0 : 1);
    prev_azimuth_ = std::atan2(y0[2], y0[1]);
  }

  void run(Outcome* out) {
    out_ = out;
    S8 d0;
    ++rhs_;
    int err = hamiltonian_rhs(st_, state_, &d0);
    if (err) return finish_failure(failure_of(err));
    update_invariants(state_);
    for (;;) {
      if (accepted_ >= pol_.max_accepted) return finish(state_, kTermStepExhaustion);
      record_step(std::fabs(step_));
      Attempt at;
      int evals = 0;
      err = attempt_step(st_, state_, step_, d0, &at, &evals);
      rhs_ += static_cast<uint64_t>(evals);
      if (err) return finish_failure(failure_of(err));
      S8 es = state_, ee = at.end;
      es[0] = time_.total();
      ee[0] = es[0] + at.dense.time_increment(1.0);
      double en = error_norm(pol_, es, ee, at.error);
      if (!std::isfinite(en)) return finish_failure(kFailNonFinite);
      if (en > 1.0) {
        ++rejected_;
        ++consecutive_rejects_;
        if (consecutive_rejects_ >= pol_.max_rejects) return finish(state_, kTermRejectExhaustion);
        if (std::fabs(step_) <= pol_.minimum_step) return finish_failure(kFailMinimumStep);
        step_ = sign_ * next_step(pol_, std::fabs(step_), en, false, step_cap(pol_, st_, state_));
        continue;
      }
      ++accepted_;
      consecutive_rejects_ = 0;
      for (double v : at.end)
        if (!std::isfinite(v)) return finish_failure(kFailNonFinite);
      Root ev{};
      ev.theta = 1.0;
      bool has_event = false;
      err = find_event(at.dense, at.end, &ev, &has_event);
      if (err) return finish_failure(failure_of(err));
      double committed_theta = has_event ? ev.theta : 1.0;
      S8 committed = has_event ? ev.state : at.end;
      record_turning_point(at.dense, at.end, has_event, ev.theta);
      record_equatorial_crossing(at.dense, at.end, has_event, ev.theta);
      time_.add(at.dense.time_increment(committed_theta));
      commit_observables(committed);
      if (has_event) {
        affine_ = std::fma(step_, ev.theta, affine_);
        return finish_event(committed, ev);
      }
      affine_ += step_;
      state_ = at.end;
      d0 = at.end_derivative;
      for (int k = 0; k < 4; ++k) {
        double v;
        if (!armed_[k] && value(k, state_, &v) == 0 && clears_band(k, v, pol_.arming_band))
          armed_[k] = true;
      }
      step_ = sign_ * next_step(pol_, std::fabs(step_), en, true, step_cap(pol_, st_, state_));
    }
  }

 private:
  bool installed(int kind) const {
    switch (kind) {
      case kSing:
        return true;
      case kHor:
        return st_.has_horizon;
      case kSurf:
        return ev_.has_surface != 0;
      default:
        return ev_.has_escape != 0;
    }
  }

  // event_value_for (geodesic.rs:678-700), horizon surface at r_+(1+offset).
  int value(int kind, const S8& y, double* v) const {
    switch (kind) {
      case kSing: {
        double m2 = st_.mass * st_.mass;
        double guard = pol_.guard_d_over_m4 * m2 * m2;
        return singularity_guard_residual(st_, &y[1], guard, v);
      }
      case kHor: {
        Geometry g;
        int e = geometry(st_, &y[1], &g);
        if (e) return e;
        double rh = st_.has_horizon ? st_.outer_horizon * (1.0 + ev_.horizon_offset) : 0.0;
        *v = g.radius - rh;
        return 0;
      }
      case kSurf:
        *v = y[3];
        return 0;
      default: {
        Geometry g;
        int e = geometry(st_, &y[1], &g);
        if (e) return e;
        *v = g.radius - (ev_.has_escape ? ev_.escape_radius : 0.0);
        return 0;
      }
    }
  }

  double scale(int kind) const {  // geodesic.rs:316-328
    switch (kind) {
      case kSing:
        return std::fmax(std::pow(st_.mass, 4), 1.0);
      case kHor:
        return std::fmax(st_.has_horizon ? st_.outer_horizon : 1.0, 1.0);
      case kSurf:
        return std::fmax(st_.mass, 1.0);
      default:
        return std::fmax(ev_.has_escape ? ev_.escape_radius : 1.0, 1.0);
    }
  }

  // localize_dense_root (geodesic.rs:640-676): bisection on the dense output.
  template <class F>
  int localize(const Dense& dn, double sv, double evv, F&& value_at, double* theta, S8* st,
               double* bracket) {
    double lo = 0.0, hi = 1.0, lv = sv, hv = evv;
    for (int it = 0; it < 64; ++it) {
      if ((hi - lo) * std::fabs(step_) <= pol_.event_affine_tol) break;
      double mid = 0.5 * (lo + hi);
      S8 ys = dn.evaluate(mid);
      for (double x : ys)
        if (!std::isfinite(x)) return kGeomNonFinite;
      double v;
      int e = value_at(ys, &v);
      if (e) return e;
      if (same_sign(lv, v)) {
        lo = mid;
        lv = v;
      } else {
        hi = mid;
        hv = v;
      }
    }
    *theta = std::fabs(lv) <= std::fabs(hv) ? lo : hi;
    *st = dn.evaluate(*theta);
    for (double x : *st)
      if (!std::isfinite(x)) return kGeomNonFinite;
    *bracket = (hi - lo) * std::fabs(step_);
    return 0;
  }

  // find_event + select_earliest_event (geodesic.rs:233-265, 722-740).
  int find_event(const Dense& dn, const S8& end, Root* out, bool* found) {
    Root roots[4];
    int n = 0;
    for (int kind = 0; kind < 4; ++kind) {
      if (!armed_[kind] || !installed(kind)) continue;
      double sv, evv;
      int e = value(kind, state_, &sv);
      if (e) return e;
      e = value(kind, end, &evv);
      if (e) return e;
      if (!crosses(kind, sv, evv)) continue;
      Root r;
      r.kind = kind;
      e = localize(
          dn, sv, evv, [&](const S8& y, double* v) { return value(kind, y, v); }, &r.theta,
          &r.state, &r.bracket);
      if (e) return e;
      double rv;
      e = value(kind, r.state, &rv);
      if (e) return e;
      r.residual = std::fabs(rv) / scale(kind);
      if (kind == kSurf) {
        Geometry g;
        if (geometry(st_, &r.state[1], &g) != 0) continue;
        if (!(g.radius >= ev_.surface_inner && g.radius <= ev_.surface_outer)) continue;
      }
      roots[n++] = r;
    }
    if (n == 0) {
      *found = false;
      return 0;
    }
    int first = 0;
    for (int i = 1; i < n; ++i)
      if (roots[i].theta < roots[first].theta) first = i;
    uint32_t cand = 0;
    for (int i = 0; i < n; ++i)
      if (std::fabs(roots[i].theta - roots[first].theta) * std::fabs(step_) <= pol_.event_tie_tol)
        cand |= 1u << roots[i].kind;
    *out = roots[first];
    candidates_ = cand;
    *found = true;
    return 0;
  }

  int traversal_radial_velocity(const S8& y, double* v) {
    ++rhs_;
    double rv;
    int e = radial_velocity(st_, y, &rv);
    if (e) return e;
    *v = sign_ * rv;
    return 0;
  }

  // record_turning_point (geodesic.rs:330-360).
  void record_turning_point(const Dense& dn, const S8& end, bool has_event, double term) {
    double s, e;
    int es = traversal_radial_velocity(state_, &s);
    int ee = traversal_radial_velocity(end, &e);
    if (es || ee) return;
    if (!((s < 0.0 && e >= 0.0) || (s > 0.0 && e <= 0.0))) return;
    double th, br;
    S8 y;
    if (localize(
            dn, s, e, [&](const S8& yy, double* v) { return traversal_radial_velocity(yy, v); },
            &th, &y, &br))
      return;
    if (has_event && !(th < term)) return;
    if (radial_turnings_ != 0xffffffffu) ++radial_turnings_;
    if (s < 0.0 && e >= 0.0) {
      Geometry g;
      if (geometry(st_, &y[1], &g) == 0) {
        turning_radius_ = has_turning_ ? std::fmin(turning_radius_, g.radius) : g.radius;
        has_turning_ = true;
      }
    }
  }

  // record_equatorial_crossing (geodesic.rs:362-383).
  void record_equatorial_crossing(const Dense& dn, const S8& end, bool has_event, double term) {
    double s = state_[3], e = end[3];
    if (!crosses(kSurf, s, e)) return;
    double th, br;
    S8 y;
    if (localize(
            dn, s, e,
            [](const S8& yy, double* v) {
              *v = yy[3];
              return 0;
            },
            &th, &y, &br))
      return;
    if (has_event && !(th < term)) return;
    if (equatorial_crossings_ != 0xffffffffu) ++equatorial_crossings_;
  }

  // commit_observables (geodesic.rs:406-418).
  void commit_observables(const S8& y) {
    double az = std::atan2(y[2], y[1]);
    double raw = az - prev_azimuth_;
    azimuth_advance_ += wrapped_angle_difference(az, prev_azimuth_);
    if (raw > M_PI) {
      if (winding_ != std::numeric_limits<int32_t>::min()) --winding_;
    } else if (raw < -M_PI) {
      if (winding_ != std::numeric_limits<int32_t>::max()) ++winding_;
    }
    prev_azimuth_ = az;
    update_invariants(y);
  }

  void update_invariants(const S8& y) {
    Invariants cur;
    if (!has_initial_ || invariants(st_, y, &cur) != 0) return;
    drift_null_ = std::fmax(drift_null_, cur.null_residual);
    drift_e_ = std::fmax(drift_e_, relative_drift(initial_.energy, cur.energy));
    drift_lz_ = std::fmax(drift_lz_, relative_drift(initial_.lz, cur.lz));
    drift_q_ = std::fmax(drift_q_, relative_drift(initial_.carter, cur.carter));
  }

  void record_step(double m) {
    min_step_ = has_steps_ ? std::fmin(min_step_, m) : m;
    max_step_ = has_steps_ ? std::fmax(max_step_, m) : m;
    has_steps_ = true;
  }

  void finish_failure(int kind) {
    finish(state_, kTermNumericalFailure);
    out_->failure = kind;
  }

  void finish_event(const S8& y, const Root& r) {
    static const int term[4] = {kTermSingularity, kTermHorizon, kTermSurface, kTermEscape};
    finish(y, term[r.kind]);
    out_->has_event = 1;
    out_->candidates = candidates_;
    out_->event_affine = affine_;
    out_->event_bracket = r.bracket;
    out_->event_residual = r.residual;
    if (r.kind == kEsc) {
      // escape_direction_xyz (geodesic.rs:480-497).
      S8 d;
      ++rhs_;
      out_->rhs_evaluations = rhs_;
      if (hamiltonian_rhs(st_, y, &d) == 0) {
        double dir[3] = {sign_ * d[1], sign_ * d[2], sign_ * d[3]};
        double n = std::sqrt(std::fma(dir[0], dir[0], std::fma(dir[1], dir[1], dir[2] * dir[2])));
        if (n > 0.0 && std::isfinite(n)) {
          out_->has_escape_direction = 1;
          for (int i = 0; i < 3; ++i) out_->escape_direction[i] = dir[i] / n;
        }
      }
    }
  }

  // finish (geodesic.rs:449-478).
  void finish(const S8& y, int termination) {
    Outcome& o = *out_;
    o = Outcome{};
    o.termination = termination;
    o.failure = kFailNone;
    o.candidates = 0;
    o.polar_side = polar_side_;
    o.radial_turnings = radial_turnings_;
    o.equatorial_crossings = equatorial_crossings_;
    o.azimuth_winding = winding_;
    o.has_turning_radius = has_turning_ ? 1 : 0;
    o.turning_radius = has_turning_ ? turning_radius_ : 0.0;
    o.azimuth_advance = azimuth_advance_;
    o.travel_time = sign_ * time_.total();
    o.affine = affine_;
    for (int i = 0; i < 8; ++i) o.state[i] = y[i];
    o.accepted = accepted_;
    o.rejected = rejected_;
    o.rhs_evaluations = rhs_;
    o.min_step = has_steps_ ? min_step_ : NAN;
    o.max_step = has_steps_ ? max_step_ : NAN;
    o.drift_null = drift_null_;
    o.drift_energy = drift_e_;
    o.drift_lz = drift_lz_;
    o.drift_carter = drift_q_;
    o.surface_status = -1;
  }

  const Spacetime st_;
  const Policy pol_;
  const Events ev_;
  S8 state_;
  double sign_;
  double step_;
  double affine_ = 0.0;
  uint64_t accepted_ = 0, rejected_ = 0, rhs_ = 0;
  uint32_t consecutive_rejects_ = 0;
  bool has_steps_ = false;
  double min_step_ = 0.0, max_step_ = 0.0;
  bool has_initial_ = false;
  Invariants initial_{};
  double drift_null_ = 0.0, drift_e_ = 0.0, drift_lz_ = 0.0, drift_q_ = 0.0;
  bool armed_[4] = {false, false, false, false};
  bool has_turning_ = false;
  double turning_radius_ = 0.0;
  uint32_t radial_turnings_ = 0, equatorial_crossings_ = 0;
  int32_t winding_ = 0;
  int polar_side_ = 1;
  double azimuth_advance_ = 0.0;
  double prev_azimuth_ = 0.0;
  CompensatedSum time_;
  uint32_t candidates_ = 0;
  Outcome* out_ = nullptr;
};

}  // namespace

// ReferencePolicy::regular_v1 / strict_v1 (policy.rs:28-66).
Policy regular_v1() {
  Policy p;
  p.pos_rtol = 2.0e-12;
  p.pos_atol = 2.0e-13;
  p.mom_rtol = 2.0e-12;
  p.mom_atol = 2.0e-13;
  p.initial_step = 1.0e-3;
  p.minimum_step = 9.094947017729282e-13;
  p.maximum_step = 0.5;
  p.safety = 0.9;
  p.min_factor = 0.2;
  p.max_factor = 5.0;
  p.max_accepted = 200000;
  p.max_rejects = 64;
  p.event_affine_tol = 2.0e-11;
  p.event_tie_tol = 5.0e-11;
  p.arming_band = 1.28e-9;
  p.guard_d_over_m4 = std::ldexp(1.0, -40);
  p.id = 0;
  p.maximum_step_per_radius = 0.0;
  return p;
}

Policy strict_v1() {
  Policy r = regular_v1();
  Policy p = r;
  p.pos_rtol = r.pos_rtol / 16.0;
  p.pos_atol = r.pos_atol / 16.0;
  p.mom_rtol = r.mom_rtol / 16.0;
  p.mom_atol = r.mom_atol / 16.0;
  p.maximum_step = 0.25;
  p.max_accepted = r.max_accepted * 2;
  p.max_rejects = r.max_rejects * 2;
  p.event_affine_tol = r.event_affine_tol / 4.0;
  p.event_tie_tol = r.event_tie_tol / 4.0;
  p.id = 1;
  return p;
}

void trace(const Spacetime& st, const Policy& pol, const Events& ev, const S8& state,
           double direction, Outcome* out) {
  Execution ex(st, pol, ev, state, direction >= 0.0 ? 1.0 : -1.0);
  ex.run(out);
}

// surface_radius_and_frequency_ratio + observable_at (surface.rs:12-185).
// Fills the surface fields of *out; returns the SurfaceError.
int observable_at(const Spacetime& st, const Emitter& em, double inner, double outer,
                  const S8& y, double omega, Outcome* out) {
  Geometry g;
  if (geometry(st, &y[1], &g)) return kSurfGeometry;
  double r = g.radius;
  if (!(r >= inner && r <= outer)) return kSurfOutside;
  double M = st.mass, a = st.spin;
  double q2 = st.charge * st.charge;
  double crs = std::fma(M, r, -q2);
  if (!std::isfinite(crs) || crs < 0.0) return kSurfOrbitUnavailable;
  double cr = std::sqrt(crs);
  double bs = a < 0.0 ? -1.0 : 1.0;
  double r2 = r * r;
  double avd = std::fma(bs * a, cr, r2);
  if (!std::isfinite(avd) || avd <= 0.0) return kSurfOrbitUnavailable;
  double omega_c = bs * cr / avd;
  double rn = std::fma(2.0 * M, r, -q2);
  double a2 = a * a;
  double tdisc = std::fma(-3.0 * M, r, std::fma(2.0 * bs * a, cr, std::fma(2.0, q2, r2)));
  if (!std::isfinite(tdisc) || tdisc <= 0.0) return kSurfNotTimelike;
  double delta = std::fma(r2, 1.0, std::fma(2.0 * M, -r, a2 + q2));
  double g_tt = -1.0 + rn / r2;
  double g_tp = -rn * a / r2;
  double rf = r2 + a2;
  double g_pp = std::fma(rf, rf, -a2 * delta) / r2;
  double cn = std::fma(omega_c, omega_c * g_pp, std::fma(2.0 * g_tp, omega_c, g_tt));
  if (!std::isfinite(cn) || cn >= 0.0) return kSurfNotTimelike;
  double ut = 1.0 / std::sqrt(-cn);
  Invariants inv;
  if (invariants(st, y, &inv)) return kSurfGeometry;
  double ef = ut * std::fma(omega_c, -inv.lz, inv.energy);
  if (!std::isfinite(ef) || ef <= 0.0) return kSurfEmitterFrequency;
  if (!std::isfinite(omega) || omega <= 0.0) return kSurfObserverFrequency;
  double ratio = omega / ef;
  if (!std::isfinite(ratio) || ratio <= 0.0) return kSurfRatio;
  // emitted_bolometric_intensity / vacuum_observed_bolometric_intensity
  // (surface.rs:167-201).
  double rr = r / M / 6.0;
  if (!std::isfinite(rr) || rr <= 0.0) return kSurfEmittedIntensity;
  double ie = em.intensity_at_six;
  for (int i = 0; i < 3; ++i) ie /= rr;
  if (!std::isfinite(ie) || ie < 0.0) return kSurfEmittedIntensity;
  double io = ie;
  for (int i = 0; i < 4; ++i) io *= ratio;
  if (!std::isfinite(io) || io < 0.0) return kSurfObservedIntensity;
  double te = 0.0, to = 0.0;
  if (em.blackbody) {
    // emitted_blackbody_temperature (surface.rs:148-165).
    double rr6 = (r / M) / 6.0;
    te = em.temperature_at_six / std::sqrt(rr6 * std::sqrt(rr6));
    if (!std::isfinite(te) || te <= 0.0) return kSurfTemperature;
    to = te * ratio;
    if (!std::isfinite(to) || to <= 0.0) return kSurfTemperature;
  }
  double cs = st.chart >= 0.0 ? a : -a;
  double az = wrapped_angle_difference(std::atan2(y[2], y[1]), std::atan2(cs, r));
  out->source_radius = r;
  out->source_azimuth = az;
  out->frequency_ratio = ratio;
  out->emitted_intensity = ie;
  out->observed_intensity = io;
  out->emitted_temperature = te;
  out->observed_temperature = to;
  return kSurfOk;
}

// Prograde equatorial ISCO of Kerr-Newman: the minimizer of the circular
// specific energy E(x) (docs/physics.md:558-566) on the timelike domain
// beyond the prograde photon orbit. E'(x) is evaluated exactly with
// forward-mode dual numbers and its sign change is bisected; the bracket is
// found by scanning outward from the photon orbit. Independent of the
// product's closed-form/Newton solver. Pinned by BPT (Q = 0) and the
// Schwarzschild 6 M / extremal RN 4 M special cases.
namespace {
struct Dual {
  double v, d;
};
Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
Dual operator/(Dual a, Dual b) { return {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
Dual dsqrt(Dual a) {
  double s = std::sqrt(a.v);
  return {s, a.d / (2.0 * s)};
}
Dual dconst(double c) { return {c, 0.0}; }
}  // namespace

double isco_prograde(double spin, double charge) {
  double a = std::fabs(spin), q2 = charge * charge;
  auto timelike = [&](double x) {
    double c = x - q2;
    if (c < 0.0) return -1.0;
    return x * x - 3.0 * x + 2.0 * q2 + 2.0 * a * std::sqrt(c);
  };
  auto energy_slope = [&](double xv) {
    Dual x{xv, 1.0};
    Dual c = dsqrt(x - dconst(q2));
    Dual num = x * x - dconst(2.0) * x + dconst(q2) + dconst(a) * c;
    Dual tl = x * x - dconst(3.0) * x + dconst(2.0 * q2) + dconst(2.0 * a) * c;
    Dual e = num / (x * dsqrt(tl));
    return e.d;
  };
  // Largest root of the timelike discriminant (prograde photon orbit): scan
  // down from large radius to the first non-timelike point, then bisect.
  double hi = 12.0, lo = hi;
  while (lo > q2) {
    lo = std::max(lo - 1e-3, q2);
    if (timelike(lo) <= 0.0) break;
  }
  for (int i = 0; i < 200; ++i) {
    double mid = 0.5 * (lo + hi);
    if (timelike(mid) > 0.0) hi = mid; else lo = mid;
  }
  double xph = hi;
  // E decreases from +inf at the photon orbit to its minimum, then rises.
  double l = xph * (1.0 + 1e-12), u = l;
  double step = 1e-3;
  while (energy_slope(u) < 0.0 && u < 1e3) {
    l = u;
    u += step;
    step *= 1.5;
  }
  for (int i = 0; i < 200; ++i) {
    double mid = 0.5 * (l + u);
    if (energy_slope(mid) < 0.0) l = mid; else u = mid;
  }
  return 0.5 * (l + u);
}

}  // namespace orc
