"""Strict reader of the reference's versioned TOML fixture documents.

Mirrors FixtureDocument::parse_toml (crates/gravlume-reference/src/fixture.rs:15-103)
and its schema parsers fixture/v1.rs (observation and geodesic documents),
fixture/v2.rs (neutral surface observable) and fixture/v3.rs (blackbody
surface transport):

* a document is UTF-8 TOML of at most 1 MiB whose `schema_version` picks the
  parser; unknown versions, profiles, fields and enum names are rejected;
* high-precision numbers travel as decimal strings, kept verbatim and
  converted to binary64 explicitly; non-finite values and negative zero are
  rejected;
* every parser accepts only the named canonical presets: after the schema
  check the document must equal, field by field and decimal string by
  decimal string, the registered preset of the same id (a changed expected
  value is a new artifact, not an edit);
* physical consistency: unit mass, the geodesic's energy at infinity and
  initial null residual, the near-critical applicability label against the
  impact parameter, the surface fixtures' base observation and their
  expected observables against the emitter law.

The results are typed and map onto this package's inputs: `scene()` gives
the `Scene` of an observation or surface fixture, `state8` the initial
state of a geodesic fixture.
"""
from __future__ import annotations

import math
import os
import re
from dataclasses import dataclass, field

import tomli

MAX_FIXTURE_BYTES = 1024 * 1024
PROFILES = {1: "baseline-v1", 2: "surface-observable-v1", 3: "surface-transport-v1"}
_DECIMAL = re.compile(r"^[+-]?(\d+(\.\d*)?|\.\d+)([eE][+-]?\d+)?$")


class FixtureError(ValueError):
    """A rejected fixture document; `kind` names the rule (FixtureError,
    fixture.rs:105-150): too-large, syntax, unsupported-schema,
    unsupported-profile, schema, preset-mismatch, invalid-physical-data,
    inconsistent-applicability, inconsistent-event-envelope,
    inconsistent-expected-observables."""

    def __init__(self, kind: str, message: str):
        super().__init__(f"{kind}: {message}")
        self.kind = kind


@dataclass(frozen=True)
class Decimal:
    """A decimal string and its binary64 value."""

    source: str
    value: float


# ---------------------------------------------------------------- schema DSL

class _Spec:
    def check(self, v, path):  # pragma: no cover - interface
        raise NotImplementedError


class _Dec(_Spec):
    def check(self, v, path):
        if not isinstance(v, str) or not _DECIMAL.match(v.strip()):
            raise FixtureError("schema", f"{path}: expected a decimal string")
        value = float(v)
        if not math.isfinite(value):
            raise FixtureError("schema", f"{path}: decimal string is not representable as f64")
        if value == 0.0 and v.strip().startswith("-"):
            raise FixtureError("schema", f"{path}: decimal string must not encode negative zero")
        return Decimal(v, value)


class _DecArray(_Spec):
    def __init__(self, n):
        self.n = n

    def check(self, v, path):
        if not isinstance(v, list) or len(v) != self.n:
            raise FixtureError("schema", f"{path}: expected {self.n} decimal strings")
        return tuple(_Dec().check(x, f"{path}[{i}]") for i, x in enumerate(v))


class _U32(_Spec):
    def check(self, v, path):
        if isinstance(v, bool) or not isinstance(v, int) or not 0 <= v < 2 ** 32:
            raise FixtureError("schema", f"{path}: expected an unsigned 32-bit integer")
        return v


class _Bool(_Spec):
    def check(self, v, path):
        if not isinstance(v, bool):
            raise FixtureError("schema", f"{path}: expected a boolean")
        return v


class _Str(_Spec):
    def check(self, v, path):
        if not isinstance(v, str):
            raise FixtureError("schema", f"{path}: expected a string")
        return v


class _Enum(_Spec):
    def __init__(self, *names):
        self.names = names

    def check(self, v, path):
        if v not in self.names:
            raise FixtureError("schema", f"{path}: unknown variant {v!r} (expected one of {self.names})")
        return v


class _Table(_Spec):
    def __init__(self, required: dict, optional: dict | None = None):
        self.required = required
        self.optional = optional or {}

    def check(self, v, path):
        if not isinstance(v, dict):
            raise FixtureError("schema", f"{path}: expected a table")
        unknown = set(v) - set(self.required) - set(self.optional)
        if unknown:
            raise FixtureError("schema", f"{path}: unknown field(s) {sorted(unknown)}")
        out = {}
        for name, spec in self.required.items():
            if name not in v:
                raise FixtureError("schema", f"{path}: missing field {name!r}")
            out[name] = spec.check(v[name], f"{path}.{name}" if path else name)
        for name, spec in self.optional.items():
            out[name] = spec.check(v[name], f"{path}.{name}" if path else name) if name in v else None
        return out


class _Tagged(_Spec):
    """An internally tagged table: `tag` picks the variant's fields."""

    def __init__(self, tag: str, variants: dict):
        self.tag = tag
        self.variants = variants

    def check(self, v, path):
        if not isinstance(v, dict) or self.tag not in v:
            raise FixtureError("schema", f"{path}: missing tag {self.tag!r}")
        name = v[self.tag]
        if name not in self.variants:
            raise FixtureError("schema", f"{path}.{self.tag}: unknown variant {name!r}")
        body = {k: x for k, x in v.items() if k != self.tag}
        out = _Table(self.variants[name]).check(body, path)
        out[self.tag] = name
        return out


D, U, B, S = _Dec(), _U32(), _Bool(), _Str()

# fixture/v1.rs:36-220
_SPACETIME = _Table({"family": _Enum("kerr-newman"), "chart": _Enum("ingoing-cartesian-kerr-schild"),
                     "signature": _Enum("mostly-plus"), "component_order": _Enum("txyz"),
                     "mass_m": D, "spin_m": D, "charge_m": D})
_V1_OBSERVATION = _Table({
    "schema_version": U, "profile": S, "id": S, "kind": _Enum("observation"),
    "evidence": _Enum("numeric", "algebraic-and-numeric"),
    "producer": _Table({"method": S, "precision_digits": U},
                       {"cross_form_azimuth_disagreement": D}),
    "spacetime": _SPACETIME,
    "observer": _Table({"coordinate_time_m": D, "oblate_radius_m": D, "polar_angle_rad": D,
                        "azimuth_rad": D, "state": _Enum("stationary"), "target_txyz_m": _DecArray(4),
                        "up_hint": _Enum("spin-positive-z"), "measured_frequency": D}),
    "viewport": _Table({"projection": _Enum("perspective"), "width": U, "height": U,
                        "vertical_fov_rad": D, "origin": _Enum("top-left"),
                        "default_subpixel": _DecArray(2)}),
    "events": _Table({"escape_radius_m": D, "singularity_guard_d_over_m4": D}),
    "expected": _Table({"parameter_state": _Enum("subextremal"), "outer_horizon_radius_m": D,
                        "observer_cartesian_xyz_m": _DecArray(3), "observer_g_tt": D}),
    "tolerance": _Table({"radius_abs_m": D, "frame_gram_max_abs": D, "initial_null_normalized_abs": D}),
})
_ORBIT = _Enum("equatorial-single-turn-scatter", "equatorial-capture")
_V1_GEODESIC = _Table({
    "schema_version": U, "profile": S, "id": S, "kind": _Enum("geodesic"),
    "evidence": _Enum("numeric", "algebraic-and-numeric"),
    "producer": _Table({"method": S, "precision_digits": U}, {"cross_form_azimuth_disagreement": D}),
    "spacetime": _SPACETIME,
    "initial": _Table({"position_txyz_m": _DecArray(4), "momentum_covariant": _DecArray(4),
                       "affine_direction": _Enum("positive", "negative"), "energy_at_infinity": D}),
    "events": _Table({}, {"escape_radius_m": D, "escape_event_initially_armed": B,
                          "outer_horizon_radius_m": D}),
    "applicability": _Tagged("class", {
        "regular": {"orbit": _ORBIT},
        "near-critical": {"orbit": _ORBIT, "critical_impact_parameter_m": D,
                          "impact_parameter_offset_m": D, "side": _Enum("escape", "capture")}}),
    "expected": _Table({"termination": _Enum("escape", "horizon-crossing"), "azimuth_advance_rad": D,
                        "initial_null_abs": D}, {"turning_radius_m": D, "event_radius_m": D}),
    "tolerance": _Table({"azimuth_advance_abs_rad": D}, {"turning_radius_abs_m": D,
                                                         "event_radius_abs_m": D}),
})
_SAMPLE = _Table({"x": U, "y": U, "subpixel": _DecArray(2)})
_SURFACE = _Table({"inner_radius_m": D, "outer_radius_m": D, "rotation": _Enum("prograde-circular")})
# fixture/v2.rs:140-245
_V2_SURFACE = _Table({
    "schema_version": U, "profile": S, "id": S, "kind": _Enum("surface-observation"),
    "evidence": _Enum("reference-convergence"), "base_observation_id": S,
    "producer": _Table({"method": S, "precision_bits": U}), "sample": _SAMPLE, "surface": _SURFACE,
    "emission": _Table({"model": _Enum("inverse-cube-bolometric-v1"), "intensity_at_6m": D}),
    "expected": _Table({"termination": _Enum("equatorial-surface"), "source_radius_m": D,
                        "source_azimuth_rad": D, "frequency_ratio": D, "travel_time_m": D,
                        "emitted_bolometric_intensity": D, "observed_bolometric_intensity": D}),
    "tolerance": _Table({"source_coordinate_abs_m": D, "source_azimuth_abs_rad": D,
                         "frequency_ratio_rel": D, "travel_time_abs_m": D,
                         "bolometric_intensity_abs": D}),
})
# fixture/v3.rs:250-376
_V3_SURFACE = _Table({
    "schema_version": U, "profile": S, "id": S, "kind": _Enum("surface-observation"),
    "evidence": _Enum("algebraic-high-precision-and-reference-convergence"),
    "base_observation_id": S, "producer": _Table({"method": S, "precision_digits": U}),
    "sample": _SAMPLE, "surface": _SURFACE,
    "emission": _Table({"model": _Enum("inverse-cube-blackbody-v1"), "intensity_at_6m": D,
                        "temperature_at_6m_kelvin": D}),
    "transport": _Tagged("model", {
        "vacuum": {},
        "pure-absorption": {"optical_depth": D},
        "constant-blackbody": {"optical_depth": D, "source_bolometric_intensity": D,
                               "source_temperature_kelvin": D},
        "pure-emission-blackbody": {"integrated_bolometric_emission": D,
                                    "emission_temperature_kelvin": D}}),
    "expected": _Table({"termination": _Enum("equatorial-surface"), "source_radius_m": D,
                        "source_azimuth_rad": D, "frequency_ratio": D, "travel_time_m": D,
                        "emitted_bolometric_intensity": D, "vacuum_observed_bolometric_intensity": D,
                        "observed_bolometric_intensity": D, "optical_depth": D,
                        "emitted_temperature_kelvin": D, "vacuum_observed_temperature_kelvin": D,
                        "observed_spectral_band_intensities": _DecArray(3)}),
    "tolerance": _Table({"source_coordinate_abs_m": D, "source_azimuth_abs_rad": D,
                         "frequency_ratio_rel": D, "travel_time_abs_m": D,
                         "bolometric_intensity_abs": D, "temperature_rel": D,
                         "spectral_intensity_abs": D}),
})


def _schema_for(version: int, kind) -> _Table:
    if version == 1:
        if kind == "observation":
            return _V1_OBSERVATION
        if kind == "geodesic":
            return _V1_GEODESIC
        raise FixtureError("schema", f"kind: unknown variant {kind!r}")
    if version == 2:
        return _V2_SURFACE
    if version == 3:
        return _V3_SURFACE
    raise FixtureError("unsupported-schema", f"schema version {version}")


# ---------------------------------------------------------------- typed results

@dataclass
class ObservationFixture:
    """A v1 observation: the validated Kerr-Schild observer and view."""

    id: str
    spin: float
    charge: float
    observer_radius: float
    polar_angle: float
    azimuth: float
    width: int
    height: int
    vertical_fov: float
    subpixel: tuple
    escape_radius: float
    expected: dict
    tolerance: dict

    def scene(self, **kw):
        """The package Scene of this observation (ingoing chart, target the
        origin, spin axis up, no surface)."""
        from .scene import INGOING, Scene
        base = dict(spin=self.spin, charge=self.charge, inclination_deg=math.degrees(self.polar_angle),
                    observer_radius=self.observer_radius, width=self.width, height=self.height,
                    fov_deg=math.degrees(self.vertical_fov), chart=INGOING, disk=None,
                    escape_radius=self.escape_radius, subpixel=self.subpixel, escape_map=False)
        base.update(kw)
        return Scene(**base)


@dataclass
class GeodesicFixture:
    """A v1 geodesic: Kerr-Newman parameters, the covariant initial state
    (t, x, y, z, p_t, p_x, p_y, p_z), the affine direction and the event
    envelope."""

    id: str
    spin: float
    charge: float
    state8: tuple
    direction: float
    escape_radius: float | None
    escape_initially_armed: bool | None
    outer_horizon_radius: float | None
    applicability: dict
    expected: dict
    tolerance: dict


@dataclass
class SurfaceObservationFixture:
    """A v2/v3 surface observation: the base observation, the sampled
    pixel, the emitter and (v3) the transport, and the expected observables."""

    id: str
    schema_version: int
    base: ObservationFixture
    sample: tuple  # (x, y, subpixel_x, subpixel_y)
    inner_radius: float
    outer_radius: float
    emission: dict
    transport: dict | None
    expected: dict
    tolerance: dict = field(default_factory=dict)

    def scene(self, **kw):
        """The package Scene: the base observation with this emitter and
        transport, sampled at the fixture's subpixel."""
        from .scene import Transport
        e = self.emission
        args = dict(disk=(self.inner_radius, self.outer_radius), intensity_at_six=e["intensity"],
                    subpixel=(self.sample[2], self.sample[3]))
        if e["model"] == "inverse-cube-blackbody-v1":
            args.update(emission="blackbody", temperature_at_six=e["temperature"])
        else:
            args.update(emission="bolometric")
        t = self.transport
        if t is not None:
            m = t["model"]
            if m == "vacuum":
                args["transport"] = Transport.vacuum()
            elif m == "pure-absorption":
                args["transport"] = Transport.pure_absorption(t["optical_depth"])
            elif m == "constant-blackbody":
                args["transport"] = Transport.constant_blackbody(
                    t["optical_depth"], t["source_bolometric_intensity"], t["source_temperature_kelvin"])
            else:
                args["transport"] = Transport.pure_emission_blackbody(
                    t["integrated_bolometric_emission"], t["emission_temperature_kelvin"])
        args.update(kw)
        return self.base.scene(**args)


# ---------------------------------------------------------------- presets

class Presets:
    """The canonical fixture artifacts a parser accepts, by id (the reference
    embeds them with include_str!, fixture/v1.rs:22-27). Built from a
    directory tree of *.toml files."""

    def __init__(self, directory: str):
        self.raw = {}
        for root, _, files in os.walk(directory):
            for name in sorted(files):
                if name.endswith(".toml"):
                    with open(os.path.join(root, name), encoding="utf-8") as f:
                        raw = _raw_document(f.read())
                    if raw["id"] in self.raw:
                        raise FixtureError("preset-mismatch", f"duplicate preset id {raw['id']!r}")
                    self.raw[raw["id"]] = raw

    def canonical(self, doc_id: str):
        return self.raw.get(doc_id)


def _raw_document(source: str) -> dict:
    if len(source.encode("utf-8")) > MAX_FIXTURE_BYTES:
        raise FixtureError("too-large", f"{len(source.encode('utf-8'))} bytes > {MAX_FIXTURE_BYTES}")
    try:
        doc = tomli.loads(source)
    except tomli.TOMLDecodeError as e:
        raise FixtureError("syntax", str(e)) from None
    version = doc.get("schema_version")
    if isinstance(version, bool) or not isinstance(version, int):
        raise FixtureError("schema", "schema_version must be an integer")
    raw = _schema_for(version, doc.get("kind")).check(doc, "")
    if raw["profile"] != PROFILES[version]:
        raise FixtureError("unsupported-profile", raw["profile"])
    return raw


def _v(d):
    return d.value if isinstance(d, Decimal) else d


def _observation(raw) -> ObservationFixture:
    sp, ob, vp = raw["spacetime"], raw["observer"], raw["viewport"]
    if sp["mass_m"].source != "1":
        raise FixtureError("preset-mismatch", "spacetime.mass_m")
    if ob["measured_frequency"].source != "1":
        raise FixtureError("invalid-physical-data", "observer frequency must be 1")
    if vp["width"] == 0 or vp["height"] == 0:
        raise FixtureError("invalid-physical-data", "empty viewport")
    return ObservationFixture(
        id=raw["id"], spin=sp["spin_m"].value, charge=sp["charge_m"].value,
        observer_radius=ob["oblate_radius_m"].value, polar_angle=ob["polar_angle_rad"].value,
        azimuth=ob["azimuth_rad"].value, width=vp["width"], height=vp["height"],
        vertical_fov=vp["vertical_fov_rad"].value,
        subpixel=tuple(x.value for x in vp["default_subpixel"]),
        escape_radius=raw["events"]["escape_radius_m"].value,
        expected={"outer_horizon_radius": raw["expected"]["outer_horizon_radius_m"].value,
                  "observer_xyz": tuple(x.value for x in raw["expected"]["observer_cartesian_xyz_m"]),
                  "observer_g_tt": raw["expected"]["observer_g_tt"].value},
        tolerance={k: v.value for k, v in raw["tolerance"].items()})


def _geodesic(raw) -> GeodesicFixture:
    sp, ini, ev = raw["spacetime"], raw["initial"], raw["events"]
    if sp["mass_m"].source != "1":
        raise FixtureError("preset-mismatch", "spacetime.mass_m")
    state = tuple(x.value for x in ini["position_txyz_m"]) + tuple(x.value for x in ini["momentum_covariant"])
    # energy at infinity exactly 1 (p_t = -1) and a null initial state (v1.rs:255-271)
    if ini["energy_at_infinity"].source != "1" or state[4] != -1.0:
        raise FixtureError("invalid-physical-data", "initial energy/null contract is inconsistent")
    null_abs = raw["expected"]["initial_null_abs"].value
    if not 0.0 <= null_abs <= 1.0e-80:
        raise FixtureError("invalid-physical-data", "initial energy/null contract is inconsistent")
    prod = raw["producer"]
    cross = prod.get("cross_form_azimuth_disagreement")
    if (raw["evidence"] != "numeric" or prod["precision_digits"] != 80 or not prod["method"].strip()
            or cross is None or not 0.0 <= cross.value <= raw["tolerance"]["azimuth_advance_abs_rad"].value):
        raise FixtureError("invalid-physical-data", "geodesic producer evidence is inconsistent")
    term = raw["expected"]["termination"]
    # event envelope (v1.rs:434-471): escape radius and its initial arming
    # together or not at all, the arming as the regular policy's 2^-12 band
    # decides it, a stored horizon radius equal to r_+ within 8 ulp
    esc, armed = ev["escape_radius_m"], ev["escape_event_initially_armed"]
    if (esc is None) != (armed is None):
        raise FixtureError("inconsistent-event-envelope", "escape radius and arming come together")
    a, q = sp["spin_m"].value, sp["charge_m"].value
    if esc is not None:
        x, y, z = state[1], state[2], state[3]
        b = x * x + y * y + z * z - a * a
        r0 = math.sqrt(0.5 * (b + math.sqrt(b * b + 4.0 * a * a * z * z)))
        if armed != (r0 - esc.value < -2.0 ** -12):
            raise FixtureError("inconsistent-event-envelope", "declared escape arming")
    hor = ev["outer_horizon_radius_m"]
    if hor is not None:
        disc = 1.0 - a * a - q * q
        if disc < 0.0 or abs(1.0 + math.sqrt(disc) - hor.value) > 8.0 * 2.220446049250313e-16:
            raise FixtureError("inconsistent-event-envelope", "stored horizon radius")
    if term == "escape" and esc is None:
        raise FixtureError("inconsistent-event-envelope", "an escape needs an escape radius")
    if term == "horizon-crossing" and hor is None:
        raise FixtureError("inconsistent-event-envelope", "a capture needs the horizon radius")
    app = raw["applicability"]
    orbit_term = {"equatorial-single-turn-scatter": "escape", "equatorial-capture": "horizon-crossing"}
    if orbit_term[app["orbit"]] != term:
        raise FixtureError("inconsistent-applicability", "orbit label against the termination")
    if app["class"] == "near-critical":
        # L_z / E of the initial state against the labelled impact parameter
        x, y, px, py = state[1], state[2], state[5], state[6]
        b = (x * py - y * px) / -state[4]
        label = app["critical_impact_parameter_m"].value + app["impact_parameter_offset_m"].value
        off = app["impact_parameter_offset_m"].value
        side_ok = (app["side"] == "escape") == (off > 0.0) and \
            (app["side"] == "escape") == (app["orbit"] == "equatorial-single-turn-scatter")
        if not side_ok or abs(b - label) > 64.0 * 2.220446049250313e-16 * max(abs(b), abs(label), 1.0):
            raise FixtureError("inconsistent-applicability", "near-critical label against the state")
    return GeodesicFixture(
        id=raw["id"], spin=sp["spin_m"].value, charge=sp["charge_m"].value, state8=state,
        direction=1.0 if ini["affine_direction"] == "positive" else -1.0,
        escape_radius=_v(ev["escape_radius_m"]), escape_initially_armed=ev["escape_event_initially_armed"],
        outer_horizon_radius=_v(ev["outer_horizon_radius_m"]),
        applicability={k: _v(v) for k, v in app.items()},
        expected={k: _v(v) for k, v in raw["expected"].items()},
        tolerance={k: _v(v) for k, v in raw["tolerance"].items()})


def _surface(raw, presets: Presets) -> SurfaceObservationFixture:
    base_raw = presets.canonical(raw["base_observation_id"])
    if base_raw is None or base_raw["kind"] != "observation":
        raise FixtureError("inconsistent-applicability", "unknown base observation")
    base = _observation(base_raw)
    smp = raw["sample"]
    sx, sy = smp["subpixel"][0].value, smp["subpixel"][1].value
    if not (smp["x"] < base.width and smp["y"] < base.height and 0.0 <= sx <= 1.0 and 0.0 <= sy <= 1.0):
        raise FixtureError("invalid-physical-data", "sample outside the viewport")
    srf, em, exp, tol = raw["surface"], raw["emission"], raw["expected"], raw["tolerance"]
    inner, outer = srf["inner_radius_m"].value, srf["outer_radius_m"].value
    if not 0.0 < inner <= outer < base.escape_radius:
        raise FixtureError("invalid-physical-data", "surface radii")
    r, g = exp["source_radius_m"].value, exp["frequency_ratio"].value
    if not (inner <= r <= outer and g > 0.0 and exp["travel_time_m"].value > 0.0):
        raise FixtureError("inconsistent-expected-observables", "source anchor outside the surface")
    # the emitter law at the expected anchor (surface.rs:167-201)
    emitted = em["intensity_at_6m"].value
    for _ in range(3):
        emitted /= r / 6.0
    observed = emitted
    for _ in range(4):
        observed *= g
    key = "vacuum_observed_bolometric_intensity" if raw["schema_version"] == 3 else \
        "observed_bolometric_intensity"
    ib = tol["bolometric_intensity_abs"].value
    if abs(emitted - exp["emitted_bolometric_intensity"].value) > ib or abs(observed - exp[key].value) > ib:
        raise FixtureError("inconsistent-expected-observables", "emitter law at the expected anchor")
    emission = {"model": em["model"], "intensity": em["intensity_at_6m"].value}
    if "temperature_at_6m_kelvin" in em:
        emission["temperature"] = em["temperature_at_6m_kelvin"].value
    transport = None
    if raw["schema_version"] == 3:
        transport = {k: _v(v) for k, v in raw["transport"].items()}
    return SurfaceObservationFixture(
        id=raw["id"], schema_version=raw["schema_version"], base=base, sample=(smp["x"], smp["y"], sx, sy),
        inner_radius=inner, outer_radius=outer, emission=emission, transport=transport,
        expected={k: (tuple(x.value for x in v) if isinstance(v, tuple) else _v(v)) for k, v in exp.items()},
        tolerance={k: v.value for k, v in tol.items()})


def parse_toml(source: str, presets: Presets):
    """Parses one fixture document against the canonical `presets`.
    Returns an ObservationFixture, GeodesicFixture or SurfaceObservationFixture;
    raises FixtureError."""
    raw = _raw_document(source)
    canonical = presets.canonical(raw["id"]) if presets is not None else None
    if canonical is None:
        raise FixtureError("preset-mismatch", f"id {raw['id']!r} is not a canonical preset")
    if raw != canonical:
        diff = sorted(k for k in raw if raw[k] != canonical.get(k))
        raise FixtureError("preset-mismatch", f"{raw['id']}: differs from the canonical artifact in {diff}")
    if raw["schema_version"] == 1:
        return _observation(raw) if raw["kind"] == "observation" else _geodesic(raw)
    return _surface(raw, presets)
