"""ctypes binding of the CPU f64 oracle (test infrastructure only).

The oracle restates Gravlume's reference CPU trace (crates/gravlume-reference,
crates/gravlume-domain) in C++ and is the checker for the HIP trace in
gravlume_amd. Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this module; the product never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")

TERM_HORIZON = 1
TERM_ESCAPE = 2
TERM_SINGULARITY = 3
TERM_STEP_EXHAUSTION = 4
TERM_NUMERICAL_FAILURE = 5
TERM_SURFACE = 7
TERM_REJECT_EXHAUSTION = 8


class Spacetime(C.Structure):
    _fields_ = [("mass", C.c_double), ("spin", C.c_double), ("charge", C.c_double),
                ("chart", C.c_double), ("extremality", C.c_int), ("has_horizon", C.c_bool),
                ("outer_horizon", C.c_double)]


class Policy(C.Structure):
    _fields_ = [("pos_rtol", C.c_double), ("pos_atol", C.c_double), ("mom_rtol", C.c_double),
                ("mom_atol", C.c_double), ("initial_step", C.c_double),
                ("minimum_step", C.c_double), ("maximum_step", C.c_double),
                ("safety", C.c_double), ("min_factor", C.c_double), ("max_factor", C.c_double),
                ("max_accepted", C.c_uint64), ("max_rejects", C.c_uint32),
                ("event_affine_tol", C.c_double), ("event_tie_tol", C.c_double),
                ("arming_band", C.c_double), ("guard_d_over_m4", C.c_double), ("id", C.c_int),
                ("maximum_step_per_radius", C.c_double)]


class Events(C.Structure):
    _fields_ = [("has_escape", C.c_int), ("escape_radius", C.c_double),
                ("has_surface", C.c_int), ("surface_inner", C.c_double),
                ("surface_outer", C.c_double), ("horizon_offset", C.c_double)]


class Emitter(C.Structure):
    _fields_ = [("intensity_at_six", C.c_double), ("blackbody", C.c_int),
                ("temperature_at_six", C.c_double)]


class Transport(C.Structure):
    _fields_ = [("slab", C.c_int), ("optical_depth", C.c_double),
                ("integrated_emission", C.c_double), ("blackbody", C.c_int),
                ("temperature", C.c_double)]


class Outcome(C.Structure):
    _fields_ = [
        ("termination", C.c_int), ("failure", C.c_int), ("candidates", C.c_uint32),
        ("polar_side", C.c_int), ("radial_turnings", C.c_uint32),
        ("equatorial_crossings", C.c_uint32), ("azimuth_winding", C.c_int32),
        ("has_turning_radius", C.c_int), ("turning_radius", C.c_double),
        ("azimuth_advance", C.c_double), ("travel_time", C.c_double), ("affine", C.c_double),
        ("state", C.c_double * 8), ("accepted", C.c_uint64), ("rejected", C.c_uint64),
        ("rhs_evaluations", C.c_uint64), ("min_step", C.c_double), ("max_step", C.c_double),
        ("drift_null", C.c_double), ("drift_energy", C.c_double), ("drift_lz", C.c_double),
        ("drift_carter", C.c_double), ("has_event", C.c_int), ("event_affine", C.c_double),
        ("event_bracket", C.c_double), ("event_residual", C.c_double),
        ("has_escape_direction", C.c_int), ("escape_direction", C.c_double * 3),
        ("surface_status", C.c_int), ("source_radius", C.c_double),
        ("source_azimuth", C.c_double), ("frequency_ratio", C.c_double),
        ("emitted_intensity", C.c_double), ("observed_intensity", C.c_double),
        ("emitted_temperature", C.c_double), ("observed_temperature", C.c_double),
    ]


class Observation(C.Structure):
    _fields_ = [
        ("st", Spacetime), ("event", C.c_double * 4), ("u", C.c_double * 4),
        ("right", C.c_double * 4), ("up", C.c_double * 4), ("arrival", C.c_double * 4),
        ("gram_residual", C.c_double), ("orientation", C.c_double),
        ("used_up_fallback", C.c_int), ("g_tt", C.c_double), ("metric", (C.c_double * 4) * 4),
        ("omega", C.c_double), ("width", C.c_uint32), ("height", C.c_uint32),
        ("fov", C.c_double), ("tan_half_fov", C.c_double),
    ]


_lib = None


def build() -> None:
    """Compile the oracle with the committed Makefile (host g++ only)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        dp = P(C.c_double)
        L.orc_trace.argtypes = [P(Spacetime), P(Policy), P(Events), dp, C.c_double, P(Outcome)]
        L.orc_trace_ordered.argtypes = [P(Spacetime), P(Policy), P(Events), dp, dp, C.c_size_t,
                                        C.c_uint, P(Outcome)]
        L.orc_observation_trace_ordered.argtypes = [P(Observation), P(Policy), P(Events),
                                                    P(Emitter), dp, C.c_size_t, C.c_uint,
                                                    P(Outcome), P(C.c_int)]
        L.orc_build_observation.argtypes = [C.c_double] * 4 + [dp, dp, dp, C.c_double, C.c_uint32,
                                                               C.c_uint32, C.c_double,
                                                               P(Observation)]
        L.orc_make_spacetime.argtypes = [C.c_double] * 4 + [P(Spacetime)]
        L.orc_geometry.argtypes = [P(Spacetime), dp, dp, dp, dp, dp, dp, dp]
        L.orc_metric.argtypes = [P(Spacetime), dp, dp, dp]
        L.orc_rhs.argtypes = [P(Spacetime), dp, dp]
        L.orc_invariants.argtypes = [P(Spacetime), dp, dp]
        L.orc_radial_velocity.argtypes = [P(Spacetime), dp, dp]
        L.orc_sight_plane.argtypes = [P(Observation), C.c_uint32, C.c_uint32, C.c_double,
                                      C.c_double, dp]
        L.orc_initial_ray.argtypes = [P(Observation), C.c_double, C.c_double, dp, dp, dp]
        L.orc_observable_at.argtypes = [P(Spacetime), P(Emitter), C.c_double, C.c_double, dp,
                                        C.c_double, P(Outcome)]
        L.orc_oblate_to_cartesian.argtypes = [C.c_double] * 5 + [dp]
        L.orc_isco_prograde.argtypes = [C.c_double, C.c_double]
        L.orc_isco_prograde.restype = C.c_double
        L.orc_compare_extremality.argtypes = [C.c_double] * 3
        L.orc_policy_preset.argtypes = [C.c_int, P(Policy)]
        L.orc_tangent_g_ordered.argtypes = ([C.c_double] * 5 + [C.c_uint32, C.c_uint32,
                                                                C.c_double, dp, C.c_size_t]
                                            + [C.c_double] * 4 + [C.c_uint, dp, dp, P(C.c_int)])
        L.orc_band_log2_fractions.argtypes = [C.c_double, dp]
        L.orc_planck_kernel_integral.argtypes = [C.c_double, C.c_double]
        L.orc_planck_kernel_integral.restype = C.c_double
        L.orc_blackbody_band_intensities.argtypes = [C.c_double, C.c_double, dp]
        L.orc_transport_bolometric.argtypes = [C.c_double, P(Transport), dp]
        L.orc_constant_source_emission.argtypes = [C.c_double, C.c_double]
        L.orc_constant_source_emission.restype = C.c_double
        L.orc_surface_radiance.argtypes = [P(Emitter), P(Transport), dp, dp, C.c_size_t, dp,
                                           P(C.c_int)]
        for name, struct in (("outcome", Outcome), ("observation", Observation),
                             ("policy", Policy), ("spacetime", Spacetime), ("events", Events),
                             ("transport", Transport), ("emitter", Emitter)):
            fn = getattr(L, f"orc_sizeof_{name}")
            fn.restype = C.c_size_t
            assert fn() == C.sizeof(struct), f"oracle struct layout mismatch: {name}"
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def policy(preset: str = "regular", **overrides) -> Policy:
    p = Policy()
    rc = lib().orc_policy_preset({"regular": 0, "strict": 1}[preset], C.byref(p))
    assert rc == 0
    for k, v in overrides.items():
        setattr(p, k, v)
    if overrides:
        p.id = 2
    return p


def scaled_policy(base: Policy, divisor: float) -> Policy:
    """The same policy with every rtol/atol divided by `divisor`."""
    p = Policy()
    C.memmove(C.byref(p), C.byref(base), C.sizeof(Policy))
    for k in ("pos_rtol", "pos_atol", "mom_rtol", "mom_atol"):
        setattr(p, k, getattr(base, k) / divisor)
    p.id = 2
    return p


def config_policy(rtol: float, atol: float, maximum_step: float = 0.5) -> Policy:
    return policy("regular", pos_rtol=rtol, pos_atol=atol, mom_rtol=rtol, mom_atol=atol,
                  maximum_step=maximum_step)


def gpu_step_policy(rtol: float, atol: float, initial_step: float, maximum_step: float,
                    maximum_step_per_radius: float, max_factor: float = 5.0) -> Policy:
    """The reference policy at the workload's tolerances with the GPU trace's
    step limits: its initial and maximum step, the relative cap
    |h| <= maximum_step_per_radius * r (not part of the reference policy) and
    the controller's growth cap max_factor (the reference's is 5); for a CPU
    baseline that takes the GPU's step sequence."""
    return policy("regular", pos_rtol=rtol, pos_atol=atol, mom_rtol=rtol, mom_atol=atol,
                  initial_step=initial_step, maximum_step=maximum_step,
                  maximum_step_per_radius=maximum_step_per_radius, max_factor=max_factor)


def spacetime(mass=1.0, spin=0.0, charge=0.0, chart=1.0) -> Spacetime:
    st = Spacetime()
    rc = lib().orc_make_spacetime(mass, spin, charge, chart, C.byref(st))
    if rc:
        raise ValueError("invalid spacetime")
    return st


def events(escape_radius=200.0, surface=None, horizon_offset=0.0, has_escape=True) -> Events:
    ev = Events()
    ev.has_escape = 1 if (has_escape and escape_radius is not None) else 0
    ev.escape_radius = escape_radius if escape_radius is not None else 0.0
    if surface is not None:
        ev.has_surface = 1
        ev.surface_inner, ev.surface_outer = surface
    ev.horizon_offset = horizon_offset
    return ev


def trace(st: Spacetime, pol: Policy, ev: Events, state, direction: float = 1.0) -> Outcome:
    y = np.ascontiguousarray(state, dtype=np.float64)
    out = Outcome()
    lib().orc_trace(C.byref(st), C.byref(pol), C.byref(ev), _dp(y), direction, C.byref(out))
    return out


def trace_ordered(st, pol, ev, states, directions, threads: int = 1):
    y = np.ascontiguousarray(states, dtype=np.float64).reshape(-1, 8)
    d = np.ascontiguousarray(directions, dtype=np.float64).reshape(-1)
    n = y.shape[0]
    out = (Outcome * n)()
    rc = lib().orc_trace_ordered(C.byref(st), C.byref(pol), C.byref(ev), _dp(y), _dp(d), n,
                                 threads, out)
    if rc:
        raise ValueError("thread count must lie in [1, 256]")
    return out


def observation(mass, spin, charge, chart, event, target, up, omega, width, height, fov):
    o = Observation()
    e = np.ascontiguousarray(event, dtype=np.float64)
    t = np.ascontiguousarray(target, dtype=np.float64)
    u = np.ascontiguousarray(up, dtype=np.float64)
    rc = lib().orc_build_observation(mass, spin, charge, chart, _dp(e), _dp(t), _dp(u), omega,
                                     width, height, fov, C.byref(o))
    if rc:
        raise ValueError(f"observation rejected (code {rc})")
    return o


def oblate_to_cartesian(spin: float, chart: float, radius: float, polar: float, azimuth: float):
    """KerrNewmanSpacetime::oblate_to_cartesian (crates/gravlume-domain/src/spacetime.rs:245-254)."""
    out = np.zeros(3)
    lib().orc_oblate_to_cartesian(spin, chart, radius, polar, azimuth, _dp(out))
    return out


def sight_plane(obs: Observation, px: int, py: int, subx=0.5, suby=0.5):
    out = np.zeros(2)
    rc = lib().orc_sight_plane(C.byref(obs), px, py, subx, suby, _dp(out))
    if rc:
        raise ValueError("sample outside the view")
    return out


def sight_planes(obs: Observation, xs, ys, subx=0.5, suby=0.5) -> np.ndarray:
    """Vectorised PerspectiveView::sight_plane for pixel centres (view.rs:107-123)."""
    xs = np.asarray(xs, dtype=np.float64)
    ys = np.asarray(ys, dtype=np.float64)
    w, h = float(obs.width), float(obs.height)
    nx = 2.0 * (xs + subx) / w - 1.0
    ny = 1.0 - 2.0 * (ys + suby) / h
    return np.stack([(w / h) * obs.tan_half_fov * nx, obs.tan_half_fov * ny], axis=-1)


def initial_ray(obs: Observation, sx: float, sy: float):
    y = np.zeros(8)
    sight = np.zeros(4)
    fn = np.zeros(2)
    rc = lib().orc_initial_ray(C.byref(obs), sx, sy, _dp(y), _dp(sight), _dp(fn))
    if rc:
        raise ValueError(f"initial ray rejected (code {rc})")
    return y, sight, fn[0], fn[1]


def observation_trace_ordered(obs, pol, ev, sight_xy, emitter: Emitter | None = None,
                              threads: int = 1):
    s = np.ascontiguousarray(sight_xy, dtype=np.float64).reshape(-1, 2)
    n = s.shape[0]
    out = (Outcome * n)()
    status = np.zeros(n, dtype=np.int32)
    em = C.byref(emitter) if emitter is not None else None
    rc = lib().orc_observation_trace_ordered(C.byref(obs), C.byref(pol), C.byref(ev), em, _dp(s),
                                             n, threads, out,
                                             status.ctypes.data_as(C.POINTER(C.c_int)))
    if rc:
        raise ValueError("thread count must lie in [1, 256]")
    return out, status


def observable_at(st, emitter, inner, outer, state, omega=1.0):
    y = np.ascontiguousarray(state, dtype=np.float64)
    out = Outcome()
    rc = lib().orc_observable_at(C.byref(st), C.byref(emitter), inner, outer, _dp(y), omega,
                                 C.byref(out))
    return rc, out


def band_log2_fractions(temperature: float):
    """log2 of the red/green/blue boxcar fractions of a blackbody
    (radiation.rs:83-101), or None when not representable."""
    out = np.zeros(3)
    if lib().orc_band_log2_fractions(float(temperature), _dp(out)):
        return None
    return out


def planck_kernel_integral(lower: float, upper: float) -> float:
    return lib().orc_planck_kernel_integral(lower, upper)


def blackbody_band_intensities(intensity: float, temperature: float):
    out = np.zeros(3)
    if lib().orc_blackbody_band_intensities(intensity, temperature, _dp(out)):
        return None
    return out


def vacuum() -> Transport:
    return Transport(0, 0.0, 0.0, 0, 0.0)


def slab(optical_depth: float, integrated_emission: float = 0.0,
         temperature: float | None = None) -> Transport:
    """HomogeneousScalarSlab with its integrated emission term (neutral when
    temperature is None)."""
    return Transport(1, optical_depth, integrated_emission, 0 if temperature is None else 1,
                     0.0 if temperature is None else temperature)


def constant_source_emission(optical_depth: float, source: float) -> float:
    return lib().orc_constant_source_emission(optical_depth, source)


def transport_bolometric(incoming: float, transport: Transport):
    out = np.zeros(3)
    if lib().orc_transport_bolometric(incoming, C.byref(transport), _dp(out)):
        return None
    return out  # (outgoing, optical depth, transmittance)


def surface_radiance(emitter: Emitter, transport: Transport, r, g):
    """The radiance chain of observable_at (surface.rs:12-60) for arrays of
    source radius r and frequency ratio g: dict of arrays plus status."""
    r = np.ascontiguousarray(r, dtype=np.float64).ravel()
    g = np.ascontiguousarray(g, dtype=np.float64).ravel()
    out = np.zeros((r.size, 9))
    status = np.zeros(r.size, dtype=np.int32)
    lib().orc_surface_radiance(C.byref(emitter), C.byref(transport), _dp(r), _dp(g), r.size,
                               _dp(out), status.ctypes.data_as(C.POINTER(C.c_int)))
    return dict(emitted=out[:, 0], vacuum=out[:, 1], observed=out[:, 2], optical_depth=out[:, 3],
                emitted_temperature=out[:, 4], observed_temperature=out[:, 5],
                bands=out[:, 6:9], status=status)


def rhs(st, state):
    y = np.ascontiguousarray(state, dtype=np.float64)
    out = np.zeros(8)
    rc = lib().orc_rhs(C.byref(st), _dp(y), _dp(out))
    if rc:
        raise ValueError(f"geometry error {rc}")
    return out


def invariants(st, state):
    y = np.ascontiguousarray(state, dtype=np.float64)
    out = np.zeros(4)
    rc = lib().orc_invariants(C.byref(st), _dp(y), _dp(out))
    if rc:
        raise ValueError(f"geometry error {rc}")
    return out


def metric(st, xyz):
    p = np.ascontiguousarray(xyz, dtype=np.float64)
    cov = np.zeros(16)
    inv = np.zeros(16)
    rc = lib().orc_metric(C.byref(st), _dp(p), _dp(cov), _dp(inv))
    if rc:
        raise ValueError(f"geometry error {rc}")
    return cov.reshape(4, 4), inv.reshape(4, 4)


def geometry(st, xyz):
    p = np.ascontiguousarray(xyz, dtype=np.float64)
    r = np.zeros(1)
    f = np.zeros(1)
    rg = np.zeros(3)
    fg = np.zeros(3)
    nc = np.zeros(4)
    nvg = np.zeros(12)
    rc = lib().orc_geometry(C.byref(st), _dp(p), _dp(r), _dp(f), _dp(rg), _dp(fg), _dp(nc),
                            _dp(nvg))
    if rc:
        raise ValueError(f"geometry error {rc}")
    return dict(radius=r[0], f=f[0], radius_gradient=rg, f_gradient=fg, null_covector=nc,
                null_vector_gradient=nvg.reshape(3, 4))


def isco_prograde(spin: float, charge: float) -> float:
    return lib().orc_isco_prograde(spin, charge)


def tangent_frequency_ratio(spin, charge, inclination, chart, observer_radius, width, height,
                            fov, sight_xy, surface, escape_radius, rtol=1e-10, threads=1):
    """Forward-mode tangent integration (tangent.cpp): g and
    dg/d(a, Q, inclination, s_x, s_y) per sight-plane point; status 0 marks
    a surface hit."""
    s = np.ascontiguousarray(sight_xy, dtype=np.float64).reshape(-1, 2)
    n = s.shape[0]
    g = np.zeros(n)
    dg = np.zeros((n, 5))
    st = np.zeros(n, dtype=np.int32)
    rc = lib().orc_tangent_g_ordered(spin, charge, inclination, chart, observer_radius, width,
                                     height, fov, _dp(s), n, surface[0], surface[1],
                                     escape_radius, rtol, threads, _dp(g), _dp(dg),
                                     st.ctypes.data_as(C.POINTER(C.c_int)))
    if rc:
        raise ValueError("thread count must lie in [1, 256]")
    return g, dg, st


def extremality(mass, spin, charge) -> int:
    return lib().orc_compare_extremality(mass, spin, charge)


def outcomes_to_arrays(outs) -> dict:
    """Structure-of-arrays view of a ctypes Outcome array (copies)."""
    n = len(outs)
    buf = np.frombuffer(outs, dtype=np.uint8).reshape(n, C.sizeof(Outcome))
    res = {}
    for name, ctype in Outcome._fields_:
        off = getattr(Outcome, name).offset
        size = C.sizeof(ctype)
        raw = np.ascontiguousarray(buf[:, off:off + size])
        if ctype in (C.c_double,):
            res[name] = raw.view(np.float64).reshape(n)
        elif ctype in (C.c_int, C.c_int32):
            res[name] = raw.view(np.int32).reshape(n)
        elif ctype is C.c_uint32:
            res[name] = raw.view(np.uint32).reshape(n)
        elif ctype is C.c_uint64:
            res[name] = raw.view(np.uint64).reshape(n)
        else:  # fixed arrays of double
            res[name] = raw.view(np.float64).reshape(n, -1)
    return res


def _wrap(d):
    return (d + np.pi) % (2 * np.pi) - np.pi


FOOTPRINT_OFFSETS = ((0.0, 0.0), (-0.25, 0.0), (0.25, 0.0), (0.0, -0.25), (0.0, 0.25))


def footprint_metrics(J):
    """Determinant, squared norm and singular values of 2x2 Jacobians
    (footprint.rs:208-238): the minor axis as |det| / major."""
    J = np.asarray(J, dtype=np.float64)
    det = J[..., 0, 0] * J[..., 1, 1] - J[..., 0, 1] * J[..., 1, 0]
    norm2 = np.sum(J * J, axis=(-2, -1))
    disc = np.sqrt(np.maximum(norm2 * norm2 - 4.0 * det * det, 0.0))
    major = np.sqrt(0.5 * (norm2 + disc))
    with np.errstate(divide="ignore", invalid="ignore"):
        minor = np.where(major == 0.0, 0.0, np.abs(det) / major)
    return det, norm2, major, minor


def surface_footprint(obs, pol, ev, emitter, px, py, subx=0.5, suby=0.5, threads=8,
                      offset=0.25):
    """ObservationTracer::surface_footprint_v1 (crates/gravlume-reference/src/
    footprint.rs:74-151): the source-chart Jacobian d(r, r_c dphi)/d(pixel x,
    y) by central differences of real traces at +-0.25 pixel, resolved only
    when the centre and all four neighbours end unambiguously on the surface
    with the same branch key. Returns dict(status (0 resolved, 1
    discontinuity, 2 centre unavailable), J (n, 2, 2), parity (1 positive,
    2 negative, 3 degenerate), singular (n, 2))."""
    px = np.atleast_1d(np.asarray(px, dtype=np.float64))
    py = np.atleast_1d(np.asarray(py, dtype=np.float64))
    n = px.size
    k = offset / 0.25
    sxy = np.stack([sight_planes(obs, px, py, subx + k * ox, suby + k * oy)
                    for ox, oy in FOOTPRINT_OFFSETS], axis=1)
    outs, st = observation_trace_ordered(obs, pol, ev, sxy.reshape(-1, 2), emitter, threads)
    A = outcomes_to_arrays(outs)
    term = A["termination"].reshape(n, 5)
    cand = A["candidates"].reshape(n, 5)
    single = (cand != 0) & ((cand & (cand - 1)) == 0)
    ok = (term == TERM_SURFACE) & (A["surface_status"].reshape(n, 5) == 0) & single & \
        (st.reshape(n, 5) == 0)
    key = np.stack([A["polar_side"], A["radial_turnings"].astype(np.int64),
                    A["equatorial_crossings"].astype(np.int64), A["azimuth_winding"]], axis=-1)
    key = key.reshape(n, 5, 4)
    same = np.all(key == key[:, :1, :], axis=(1, 2))
    r = A["source_radius"].reshape(n, 5)
    phi = A["source_azimuth"].reshape(n, 5)
    arc = r[:, :1] * _wrap(phi - phi[:, :1])
    J = np.zeros((n, 2, 2))
    span = 2.0 * offset
    J[:, 0, 0] = (r[:, 2] - r[:, 1]) / span
    J[:, 0, 1] = (r[:, 4] - r[:, 3]) / span
    J[:, 1, 0] = (arc[:, 2] - arc[:, 1]) / span
    J[:, 1, 1] = (arc[:, 4] - arc[:, 3]) / span
    status = np.where(~ok[:, 0], 2, np.where(np.all(ok, axis=1) & same, 0, 1))
    J[status != 0] = 0.0
    det, norm2, major, minor = footprint_metrics(J)
    floor = 64.0 * np.finfo(np.float64).eps * norm2
    parity = np.where(det > floor, 1, np.where(det < -floor, 2, 3))
    return dict(status=status, J=J, parity=parity, singular=np.stack([major, minor], axis=1))

