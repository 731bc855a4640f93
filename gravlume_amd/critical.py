"""Kerr-Newman critical curve and near-critical initial data, in f64 (M = 1).

Spherical photon orbits of radius r have the constants of motion (E = 1)
    xi(r)  = (r^2 + a^2 - 4 r Delta / Delta') / a,
    eta(r) = 16 r^2 Delta / Delta'^2 - (xi - a)^2,
with Delta = r^2 - 2 r + a^2 + Q^2 and Delta' = 2 r - 2: the double roots of
the radial potential R(r) = (r^2 + a^2 - a xi)^2 - Delta (eta + (xi - a)^2).
The orbits with eta >= 0 lie between the prograde and retrograde circular
photon orbits r_1 < r_2 (eta(r_1) = eta(r_2) = 0). Seen by a distant observer
at inclination theta they trace the critical curve (the shadow edge) in the
celestial coordinates alpha = -xi / sin(theta), beta = +-sqrt(Theta), with
Theta = eta + a^2 cos^2(theta) - xi^2 cot^2(theta).

near_critical_states() places rays at a signed relative offset delta from
that curve along the ray from its centre at a uniformly drawn position angle
and returns their exact f64 initial states (E, L_z = xi, Carter = eta) at a
static observer at oblate (r_obs, theta, azimuth 0) in Cartesian Kerr-Schild
coordinates, ready for GeodesicBatch.trace_ordered (backward direction).
For a -> 0 the orbits collapse onto the photon sphere and the curve is the
circle b^2 = xi^2 + eta = r^4 / Delta(r_ph).
"""
from __future__ import annotations

import math

import numpy as np


def _delta(a, q, r):
    return r * r - 2.0 * r + a * a + q * q


def photon_orbit_constants(a: float, q: float, r):
    """(xi, eta) of the spherical photon orbit of radius r (a != 0)."""
    r = np.asarray(r, dtype=np.float64)
    d = _delta(a, q, r)
    dp = 2.0 * r - 2.0
    xi = (r * r + a * a - 4.0 * r * d / dp) / a
    eta = 16.0 * r * r * d / (dp * dp) - (xi - a) ** 2
    return xi, eta


def photon_sphere(q: float) -> float:
    """Reissner-Nordstrom photon-sphere radius (3 + sqrt(9 - 8 Q^2)) / 2."""
    return 0.5 * (3.0 + math.sqrt(9.0 - 8.0 * q * q))


def circular_photon_orbits(a: float, q: float) -> tuple[float, float]:
    """Prograde and retrograde equatorial photon orbits r_1 < r_2 (roots of
    eta(r) = 0 on either side of the photon sphere), by bisection."""
    rps = photon_sphere(q)
    r_plus = 1.0 + math.sqrt(max(0.0, 1.0 - a * a - q * q))

    def eta(r):
        return float(photon_orbit_constants(a, q, r)[1])

    def root(lo, hi):
        flo = eta(lo)
        for _ in range(200):
            mid = 0.5 * (lo + hi)
            fm = eta(mid)
            if (fm > 0.0) == (flo > 0.0):
                lo, flo = mid, fm
            else:
                hi = mid
        return 0.5 * (lo + hi)

    lo = max(r_plus, 1.0) * (1.0 + 1e-12) + 1e-12
    return root(lo, rps), root(rps, 4.0 + 1e-9)


def critical_curve(a: float, q: float, theta: float, samples: int = 4096):
    """Celestial (alpha, beta) of the upper half (beta >= 0) of the critical
    curve for an observer at inclination theta, with the orbit radii."""
    if abs(a) < 1e-9:
        rps = photon_sphere(q)
        b = rps * rps / math.sqrt(_delta(0.0, q, rps))
        psi = np.linspace(0.0, math.pi, samples)
        return b * np.cos(psi), b * np.sin(psi), np.full(samples, rps)
    r1, r2 = circular_photon_orbits(a, q)
    r = np.linspace(r1, r2, samples)
    xi, eta = photon_orbit_constants(a, q, r)
    ct, st = math.cos(theta), math.sin(theta)
    big_theta = eta + a * a * ct * ct - xi * xi * (ct / st) ** 2
    keep = big_theta >= 0.0
    return -xi[keep] / st, np.sqrt(big_theta[keep]), r[keep]


def _state_at_observer(a, q, chart, r_obs, theta, xi, eta, p_theta):
    """Cartesian Kerr-Schild initial state (t, x, y, z, p_t, p_x, p_y, p_z) of
    the photon with E = 1, L_z = xi, Carter eta, polar momentum p_theta, at
    oblate (r_obs, theta, azimuth 0), moving outward in the future."""
    s = 1.0 if chart >= 0 else -1.0
    r, th = r_obs, theta
    d = _delta(a, q, r)
    big_r = (r * r + a * a - a * xi) ** 2 - d * (eta + (xi - a) ** 2)
    p_r_bl = np.sqrt(np.maximum(big_r, 0.0)) / d
    # Kerr-Schild chart: t_KS = t + s (T(r) - r), phi_KS = phi + s Phi(r),
    # T' = (r^2 + a^2) / Delta, Phi' = a / Delta; p_t = -1, p_phi = xi
    p_r = p_r_bl + s * ((r * r + a * a) / d - 1.0) - s * a * xi / d
    p_phi = xi
    # oblate (r, theta, phi) -> Cartesian: x + i y = (r + i s a) e^{i phi} sin(theta), z = r cos(theta)
    sth, cth = math.sin(th), math.cos(th)
    x, y, z = r * sth, s * a * sth, r * cth
    # Jacobian d(x, y, z)/d(r, theta, phi) at phi = 0
    J = np.array([[sth, r * cth, -s * a * sth],
                  [0.0, s * a * cth, r * sth],
                  [cth, -r * sth, 0.0]])
    Jinv_T = np.linalg.inv(J).T
    p_sph = np.stack([p_r, p_theta, p_phi * np.ones_like(p_r)], axis=-1)
    p_xyz = p_sph @ Jinv_T.T
    n = p_xyz.shape[0]
    out = np.zeros((n, 8))
    out[:, 1], out[:, 2], out[:, 3] = x, y, z
    out[:, 4] = -1.0
    out[:, 5:8] = p_xyz
    return out


def _celestial(a, q, theta, r):
    """(alpha, beta >= 0, Theta) of the photon orbit of radius r."""
    xi, eta = photon_orbit_constants(a, q, r)
    ct, st = math.cos(theta), math.sin(theta)
    big_theta = eta + a * a * ct * ct - xi * xi * (ct / st) ** 2
    return -xi / st, np.sqrt(np.maximum(big_theta, 0.0)), big_theta


def _bisect(f, lo, hi, iters=200):
    """Vectorized bisection for a sign change of f on [lo, hi] (arrays)."""
    lo = np.array(lo, dtype=np.float64)
    hi = np.array(hi, dtype=np.float64)
    flo = f(lo)
    for _ in range(iters):
        mid = 0.5 * (lo + hi)
        fm = f(mid)
        same = (fm > 0.0) == (flo > 0.0)
        lo = np.where(same, mid, lo)
        flo = np.where(same, fm, flo)
        hi = np.where(same, hi, mid)
    return 0.5 * (lo + hi)


def near_critical_states(a: float, q: float, r_obs: float, theta: float, count: int,
                         seed: int, chart: int = -1, delta_range=(1e-8, 1e-3),
                         uniform_angle: bool = True):
    """`count` rays at signed offsets delta ~ +-log-U[delta_range] from the
    critical curve at uniformly drawn position angles psi about the curve's
    centre: the curve point at psi is solved exactly (bisection on the orbit
    radius), then moved radially by the factor 1 + delta. Returns
    (states (count, 8), delta (count,), xi, eta): delta > 0 lies outside the
    shadow edge (the ray escapes after orbiting), delta < 0 inside. With
    uniform_angle=False the curve points are drawn uniformly in orbit radius
    instead (no root solve; faster for large counts)."""
    if abs(a) < 1e-9:
        raise ValueError("near-critical data needs a != 0 (use the photon sphere for a = 0)")
    rng = np.random.default_rng(seed)
    r1, r2 = circular_photon_orbits(a, q)
    # visible orbit radii: Theta(theta) >= 0 (a sub-interval of [r1, r2])
    grid = np.linspace(r1, r2, 8193)
    _, _, tg = _celestial(a, q, theta, grid)
    vis = np.flatnonzero(tg >= 0.0)
    lo_i, hi_i = vis[0], vis[-1]
    theta_f = lambda r: _celestial(a, q, theta, r)[2]  # noqa: E731
    r_lo = grid[lo_i] if lo_i == 0 else float(_bisect(theta_f, grid[lo_i - 1], grid[lo_i]))
    r_hi = grid[hi_i] if hi_i == grid.size - 1 else float(_bisect(theta_f, grid[hi_i], grid[hi_i + 1]))
    al, _, _ = _celestial(a, q, theta, np.array([r_lo, r_hi]))
    centre = 0.5 * (al[0] + al[1])

    def psi_of(r):
        al_, be_, _ = _celestial(a, q, theta, r)
        return np.arctan2(be_, al_ - centre)

    psi_lo, psi_hi = psi_of(np.array([r_lo]))[0], psi_of(np.array([r_hi]))[0]
    psi = rng.uniform(-math.pi, math.pi, count)
    if uniform_angle:
        upper = np.clip(np.abs(psi), min(psi_lo, psi_hi), max(psi_lo, psi_hi))
        rp = _bisect(lambda r: psi_of(r) - upper, np.full(count, r_lo), np.full(count, r_hi),
                     iters=64)
    else:
        rp = rng.uniform(r_lo, r_hi, count)
    alpha_c, beta_c, _ = _celestial(a, q, theta, rp)
    lo, hi = math.log(delta_range[0]), math.log(delta_range[1])
    delta = np.exp(rng.uniform(lo, hi, count)) * np.where(rng.uniform(size=count) < 0.5, -1.0, 1.0)
    alpha = centre + (1.0 + delta) * (alpha_c - centre)
    beta = (1.0 + delta) * beta_c * np.where(psi < 0.0, -1.0, 1.0)
    st, ct = math.sin(theta), math.cos(theta)
    xi = -alpha * st
    eta = beta * beta - a * a * ct * ct + xi * xi * (ct / st) ** 2
    # beta is the polar momentum at infinity; at the observer Theta(theta)
    # has the same value, and its sign picks the half of the curve
    states = _state_at_observer(a, q, chart, r_obs, theta, xi, eta, beta)
    return states, delta, xi, eta
