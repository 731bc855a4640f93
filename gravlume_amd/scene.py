"""Scene descriptions: validated Kerr-Newman observations for the GPU trace.

Mirrors the reference's domain seam (crates/gravlume-domain/src/scene.rs:99-143,
observer.rs:9-104, view.rs:28-69): a spacetime (M = 1, spin a, charge Q,
Kerr-Schild chart), a stationary observer at an oblate position looking at
the origin with spin-axis up, a pinhole viewport, the event surfaces
(escape radius, horizon capture surface, prograde equatorial emitter) and the
adaptive DP5(4) policy.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _capi

INGOING = 1
OUTGOING = -1


def oblate_to_cartesian(spin: float, chart: int, radius: float, polar: float,
                        azimuth: float = 0.0) -> tuple[float, float, float]:
    """Chart oblate coordinates to Cartesian Kerr-Schild (spacetime.rs:245-254)."""
    st, ct = math.sin(polar), math.cos(polar)
    sp, cp = math.sin(azimuth), math.cos(azimuth)
    cs = chart * spin
    return ((radius * cp - cs * sp) * st, (radius * sp + cs * cp) * st, radius * ct)


def outer_horizon(spin: float, charge: float) -> float:
    d = 1.0 - spin * spin - charge * charge
    if d < 0.0:
        raise ValueError("superextremal spacetime has no horizon")
    return 1.0 + math.sqrt(d)


def isco_prograde(spin: float, charge: float = 0.0) -> float:
    """Prograde equatorial ISCO radius of Kerr-Newman (M = 1).

    Q = 0: the Bardeen-Press-Teukolsky closed form. Q != 0: the root of the
    analytic dE/dx for the circular specific energy
    E(x) = (x^2 - 2x + Q^2 + a sqrt(x - Q^2)) / (x sqrt(x^2 - 3x + 2Q^2 + 2a sqrt(x - Q^2)))
    (the reference's docs/physics.md:558-566), bracketed from outside and
    bisected.
    """
    a = abs(spin)
    if a > 1.0:
        raise ValueError("|a| > 1")
    z1 = 1.0 + (1.0 - a * a) ** (1.0 / 3.0) * ((1.0 + a) ** (1.0 / 3.0) + (1.0 - a) ** (1.0 / 3.0))
    z2 = math.sqrt(3.0 * a * a + z1 * z1)
    r_bpt = 3.0 + z2 - math.sqrt((3.0 - z1) * (3.0 + z1 + 2.0 * z2))
    q2 = charge * charge
    if q2 == 0.0:
        return r_bpt

    def slope(x: float) -> float:
        # analytic dE/dx
        s = math.sqrt(x - q2)
        num = x * x - 2 * x + q2 + a * s
        dnum = 2 * x - 2 + a / (2 * s)
        tl = x * x - 3 * x + 2 * q2 + 2 * a * s
        dtl = 2 * x - 3 + a / s
        den = x * math.sqrt(tl)
        dden = math.sqrt(tl) + x * dtl / (2 * math.sqrt(tl))
        return (dnum * den - num * dden) / (den * den)

    def timelike(x: float) -> bool:
        return x > q2 and x * x - 3 * x + 2 * q2 + 2 * a * math.sqrt(x - q2) > 0.0

    # Bracket the minimum of E between the photon orbit (where E -> inf and
    # dE/dx < 0) and the first radius with dE/dx > 0, then bisect on the sign
    # of the analytic slope.
    hi = max(r_bpt, 6.0) * 2.0
    while slope(hi) <= 0.0:
        hi *= 2.0
    lo = hi
    while timelike(lo * 0.99) and slope(lo * 0.99) > 0.0:
        lo *= 0.99
    lo *= 0.99
    while not timelike(lo):
        lo = 0.5 * (lo + hi)
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if mid == lo or mid == hi:
            break
        if slope(mid) < 0.0:
            lo = mid
        else:
            hi = mid
    x = 0.5 * (lo + hi)
    return x


@dataclass(frozen=True)
class Transport:
    """Transport between the equatorial source and the observer
    (SurfaceTransport, crates/gravlume-domain/src/scene.rs:40-43): vacuum,
    or a homogeneous grey slab that maps I_in to
    I_in exp(-optical_depth) + integrated_emission, the emission term being a
    blackbody at `temperature` K, or neutral when temperature is None
    (HomogeneousScalarSlab, radiation.rs:54-220). The constructors below
    follow the reference's; the values are checked at gl_trace_plan_create."""

    slab: bool = False
    optical_depth: float = 0.0
    integrated_emission: float = 0.0
    temperature: float | None = None

    @staticmethod
    def vacuum() -> "Transport":
        return Transport()

    @staticmethod
    def pure_absorption(optical_depth: float) -> "Transport":
        return Transport(True, optical_depth, 0.0, None)

    @staticmethod
    def constant_bolometric(optical_depth: float, source: float) -> "Transport":
        """Constant neutral source function: emission S (1 - e^-tau)."""
        return Transport(True, optical_depth, source * -math.expm1(-optical_depth), None)

    @staticmethod
    def constant_blackbody(optical_depth: float, source: float, temperature: float) -> "Transport":
        return Transport(True, optical_depth, source * -math.expm1(-optical_depth), temperature)

    @staticmethod
    def pure_emission_bolometric(emission: float) -> "Transport":
        return Transport(True, 0.0, emission, None)

    @staticmethod
    def pure_emission_blackbody(emission: float, temperature: float) -> "Transport":
        return Transport(True, 0.0, emission, temperature)


@dataclass
class Scene:
    """One image scene; `to_desc()` packs it for gl_trace_plan_create.

    The equatorial emitter radiates I(r) = intensity_at_six (r / 6)^-3
    (emission.rs:12-145); with emission == "blackbody" as a diluted
    blackbody at T(r) = T_6 (r / 6)^-3/4, T_6 = temperature_at_six, or, when
    that is None, the T_6 that puts temperature_inner at the inner radius."""
    spin: float = 0.94
    charge: float = 0.0
    inclination_deg: float = 80.0
    observer_radius: float = 100.0
    width: int = 1920
    height: int = 1080
    fov_deg: float = 15.0
    chart: int = OUTGOING
    disk: tuple[float, float] | None = None  # (inner, outer) in M; inner None -> ISCO
    emission: str = "blackbody"  # "bolometric" or "blackbody"
    intensity_at_six: float = 1.0
    temperature_at_six: float | None = None
    temperature_inner: float = 1.0e7  # K at the inner radius, when temperature_at_six is None
    transport: Transport = field(default_factory=Transport)
    escape_radius: float = 1000.0
    horizon_offset: float = 1.0e-6
    rtol: float = 1.0e-6
    atol: float = 1.0e-8
    initial_step: float = 0.05  # first step as a fraction of the observer radius
    minimum_step: float = 2.0 ** -30
    maximum_step: float = 256.0
    maximum_step_per_radius: float = 0.25
    step_safety: float = 0.9       # controller safety factor (policy.rs:160-173)
    step_min_factor: float = 0.2
    # growth cap per accepted step: 2 (the reference policy's is 5). Smaller
    # growth keeps the 64 rays of a wavefront nearer in step count, so fewer
    # lanes idle while their neighbours finish: on the 1920x1080 a = 0.94
    # frame 4.3% fewer wave instructions for 0.9% more attempts
    step_max_factor: float = 2.0
    max_accepted_steps: int = 100_000
    max_consecutive_rejects: int = 64
    event_tie_tolerance: float = 2.0 ** -17
    arming_band: float = 2.0 ** -12
    drift_limit: float = 0.05
    reproject_interval: int = 16
    subpixel: tuple[float, float] = (0.5, 0.5)
    options: int = (_capi.OPT_F64_DISK_ROOT | _capi.OPT_F64_HORIZON_STEPS)
    capture_certificate: bool = False  # interval radial-capture certificate
    # near-critical test (0: the library's 16; negative: off) and the f64
    # refine of near-critical rays at refine_rtol / refine_atol (0: 1e-12 / 1e-14)
    near_critical_factor: float = 0.0
    branch_factor: float = 0.0  # 0: the library's 4096
    refine: bool = False
    refine_rtol: float = 0.0
    refine_atol: float = 0.0
    escape_map: bool = True
    extra: dict = field(default_factory=dict)

    @property
    def rays(self) -> int:
        return self.width * self.height

    def disk_radii(self) -> tuple[float, float] | None:
        if self.disk is None:
            return None
        inner, outer = self.disk
        if inner is None:
            inner = isco_prograde(self.spin, self.charge)
        return (inner, outer)

    def six_temperature(self) -> float:
        """T_6 of the blackbody emitter (0 for a bolometric one)."""
        if self.emission != "blackbody":
            return 0.0
        if self.temperature_at_six is not None:
            return self.temperature_at_six
        inner = self.disk_radii()[0]
        return self.temperature_inner * (inner / 6.0) ** 0.75

    def observer_xyz(self) -> tuple[float, float, float]:
        return oblate_to_cartesian(self.spin, self.chart, self.observer_radius,
                                   math.radians(self.inclination_deg), 0.0)

    def to_desc(self) -> _capi.SceneDesc:
        d = _capi.SceneDesc()
        d.mass = 1.0
        d.spin = self.spin
        d.charge = self.charge
        d.chart = self.chart
        opts = self.options
        if self.capture_certificate:
            opts |= _capi.OPT_CAPTURE_CERTIFICATE
        if self.escape_map:
            opts |= _capi.OPT_ESCAPE_MAP
        if self.refine:
            opts |= _capi.OPT_NEAR_CRITICAL_REFINE
        d.options = opts
        x, y, z = self.observer_xyz()
        d.observer_txyz[:] = [0.0, x, y, z]
        d.target_txyz[:] = [0.0, 0.0, 0.0, 0.0]
        d.up_hint[:] = [0.0, 0.0, 1.0]
        d.frequency = 1.0
        d.width, d.height = self.width, self.height
        d.vertical_fov = math.radians(self.fov_deg)
        d.subpixel[:] = list(self.subpixel)
        d.escape_radius = self.escape_radius
        d.horizon_offset = self.horizon_offset
        radii = self.disk_radii()
        if radii is not None:
            d.has_surface = 1
            d.surface_inner, d.surface_outer = radii
            if self.emission not in ("bolometric", "blackbody"):
                raise ValueError(f"unknown emission model {self.emission!r}")
            d.emission_model = (_capi.EMISSION_BLACKBODY if self.emission == "blackbody"
                                else _capi.EMISSION_BOLOMETRIC)
            d.intensity_at_six = self.intensity_at_six
            d.temperature_at_six = self.six_temperature()
            tr = self.transport
            d.transport_model = _capi.TRANSPORT_SLAB if tr.slab else _capi.TRANSPORT_VACUUM
            d.optical_depth = tr.optical_depth
            d.slab_emission = tr.integrated_emission
            d.slab_temperature = 0.0 if tr.temperature is None else tr.temperature
        d.rtol, d.atol = self.rtol, self.atol
        d.initial_step = self.initial_step * self.observer_radius
        d.minimum_step = self.minimum_step
        d.maximum_step = self.maximum_step
        d.maximum_step_per_radius = self.maximum_step_per_radius
        d.step_safety = self.step_safety
        d.step_min_factor = self.step_min_factor
        d.step_max_factor = self.step_max_factor
        d.max_accepted_steps = self.max_accepted_steps
        d.max_consecutive_rejects = self.max_consecutive_rejects
        d.event_tie_tolerance = self.event_tie_tolerance
        d.arming_band = self.arming_band
        d.drift_limit = self.drift_limit
        d.reproject_interval = self.reproject_interval
        d.near_critical_factor = self.near_critical_factor
        d.branch_factor = self.branch_factor
        d.refine_rtol = self.refine_rtol
        d.refine_atol = self.refine_atol
        return d


def escape_map(seed: int = 20261203, width: int = 2048, height: int = 1024) -> np.ndarray:
    """Seeded uniform-random RGB equirectangular map, shape (3, height, width), f32."""
    rng = np.random.default_rng(seed)
    return rng.random((3, height, width), dtype=np.float32)


# The measured workloads (Kerr-Newman, M = 1).
def single_frame_1080p(**kw) -> Scene:
    """1920x1080, a = 0.94, Q = 0, observer r = 100 M at 80 deg, FOV 15 deg,
    thin prograde disk [r_ISCO, 30 M], escape to r = 1000 M into the map."""
    base = dict(spin=0.94, charge=0.0, inclination_deg=80.0, observer_radius=100.0,
                width=1920, height=1080, fov_deg=15.0, disk=(None, 30.0))
    base.update(kw)
    return Scene(**base)


def small_cpu_scene(**kw) -> Scene:
    """64x64, a = 0.5, Q = 0.3, observer r = 50 M at 60 deg, FOV 20 deg,
    disk [r_ISCO, 20 M]."""
    base = dict(spin=0.5, charge=0.3, inclination_deg=60.0, observer_radius=50.0, width=64,
                height=64, fov_deg=20.0, disk=(None, 20.0))
    base.update(kw)
    return Scene(**base)


def adjoint_scene(**kw) -> Scene:
    """512x512, a = 0.9, Q = 0.2, observer r = 100 M at 70 deg, FOV 20 deg,
    thin disk [r_ISCO, 25 M]; the scene of the adjoint (VJP) workload."""
    base = dict(spin=0.9, charge=0.2, inclination_deg=70.0, observer_radius=100.0, width=512,
                height=512, fov_deg=20.0, disk=(None, 25.0), escape_map=False)
    base.update(kw)
    return Scene(**base)


def near_critical_scene(**kw) -> Scene:
    """Kerr-Newman a = 0.998, Q = 0.05, observer r = 200 M at 85 deg, no disk,
    escape at 1000 M: the spacetime and policy of the near-critical batch
    (initial states from gravlume_amd.critical)."""
    base = dict(spin=0.998, charge=0.05, inclination_deg=85.0, observer_radius=200.0,
                width=64, height=64, fov_deg=15.0, disk=None, escape_map=False)
    base.update(kw)
    return Scene(**base)


def uneven_scene_draws(seed: int = 20261206, **kw):
    """The stream of seed-drawn scenes of uneven size: a ~ U[0, 0.998],
    Q ~ U[0, 0.9 sqrt(1 - a^2)], inclination ~ U[5, 89] deg,
    r_obs ~ U[30, 200] M, side in {128..1024}, FOV ~ U[8, 30] deg; disk
    [r_ISCO, 20 M], escape at 1000 M. An endless generator of candidates."""
    rng = np.random.default_rng(seed)
    sides = [128, 256, 512, 768, 1024]
    while True:
        a = float(rng.uniform(0.0, 0.998))
        q = float(rng.uniform(0.0, math.sqrt(1.0 - a * a) * 0.9))
        side = int(rng.choice(sides))
        s = Scene(spin=a, charge=q, inclination_deg=float(rng.uniform(5.0, 89.0)),
                  observer_radius=float(rng.uniform(30.0, 200.0)), width=side, height=side,
                  fov_deg=float(rng.uniform(8.0, 30.0)), disk=(None, 20.0), escape_map=False)
        for k, v in kw.items():
            setattr(s, k, v)
        yield s


def uneven_scenes(count: int = 256, seed: int = 20261206, **kw) -> list[Scene]:
    """The first `count` draws of uneven_scene_draws that pass the scene
    checks (gl_scene_validate, which accepts exactly the observers the
    reference's frame construction accepts). For the default seed no draw
    of the first 256 is refused, so this is the plain draw sequence."""
    lib = _capi.load()
    out = []
    for s in uneven_scene_draws(seed, **kw):
        if len(out) >= count:
            break
        d = s.to_desc()
        if lib.gl_scene_validate(C.byref(d)) != 0:
            continue
        out.append(s)
    return out


def shadow_edge_requests(scene: Scene, fraction: float, seed: int,
                         delta_range=(1e-6, 1e-2)) -> tuple[np.ndarray, np.ndarray]:
    """The jittered share of a scene's rays as geodesic-batch requests: a
    `fraction` of its pixel count, each a backward ray from the scene's
    observer at a signed offset delta ~ +-log-U[delta_range] from the
    critical curve (the shadow edge), so its step count runs from tens to
    thousands. Returns (states (n, 8) f64, delta (n,))."""
    from .critical import near_critical_states
    n = int(round(fraction * scene.rays))
    if n == 0:
        return np.zeros((0, 8)), np.zeros(0)
    states, delta, _, _ = near_critical_states(scene.spin, scene.charge, scene.observer_radius,
                                               math.radians(scene.inclination_deg), n, seed,
                                               chart=scene.chart, delta_range=delta_range,
                                               uniform_angle=False)
    return states, delta
