"""Blackbody bands and surface transport: the oracle (radiation.cpp) against
the reference's surface-transport fixtures, known answers and an independent
high-precision quadrature; the product's host band table against the oracle;
and the scene checks of the emitter and slab at the C ABI (no GPU)."""
import ctypes as C
import json
import math
import os

import mpmath
import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = json.load(open(os.path.join(HERE, "golden", "surface_transport_fixtures.json")))
CASES = sorted(k for k in FIX if not k.startswith("_"))

# The reference's error budgets for the band table and the GPU radiance
# (crates/gravlume-render/src/scientific_capture.rs:8-12).
LUT_ABS_FRACTION = 3.0e-6
LUT_VISIBLE_REL = 2.0e-3
LUT_REL_FLOOR = 1.0e-6


def _transport(O, t):
    m = t["model"]
    if m == "vacuum":
        return O.vacuum()
    if m == "pure-absorption":
        return O.slab(float(t["optical_depth"]))
    if m == "constant-blackbody":
        tau = float(t["optical_depth"])
        return O.slab(tau, O.constant_source_emission(tau, float(t["source_bolometric_intensity"])),
                      float(t["source_temperature_kelvin"]))
    if m == "pure-emission-blackbody":
        return O.slab(0.0, float(t["integrated_bolometric_emission"]),
                      float(t["emission_temperature_kelvin"]))
    raise AssertionError(m)


@pytest.mark.parametrize("case", CASES)
def test_oracle_reproduces_the_surface_transport_fixture(oracle, case):
    """fixtures/v3: from the fixture's own (r, g) the oracle's emitted and
    observed bolometric intensities, temperatures and transported band
    intensities match the 80-digit values within the fixture tolerances."""
    f = FIX[case]
    e, tol = f["expected"], f["tolerance"]
    em = oracle.Emitter(float(f["emission"]["intensity_at_6m"]), 1,
                        float(f["emission"]["temperature_at_6m_kelvin"]))
    s = oracle.surface_radiance(em, _transport(oracle, f["transport"]),
                                [float(e["source_radius_m"])], [float(e["frequency_ratio"])])
    assert s["status"][0] == 0
    ib = float(tol["bolometric_intensity_abs"])
    assert abs(s["emitted"][0] - float(e["emitted_bolometric_intensity"])) <= ib
    assert abs(s["vacuum"][0] - float(e["vacuum_observed_bolometric_intensity"])) <= ib
    assert abs(s["observed"][0] - float(e["observed_bolometric_intensity"])) <= ib
    assert abs(s["optical_depth"][0] - float(e["optical_depth"])) <= 1e-15
    tr = float(tol["temperature_rel"])
    assert abs(s["emitted_temperature"][0] / float(e["emitted_temperature_kelvin"]) - 1) <= tr
    assert abs(s["observed_temperature"][0] / float(e["vacuum_observed_temperature_kelvin"]) - 1) <= tr
    want = np.array([float(v) for v in e["observed_spectral_band_intensities"]])
    assert np.max(np.abs(s["bands"][0] - want)) <= float(tol["spectral_intensity_abs"])


def test_oracle_band_fractions_known_answers(oracle):
    """The reference's independent 6000 K oracle and its diluted 220 K case
    (radiation.rs:249-287, spectral_lut.rs:156-169)."""
    got = oracle.blackbody_band_intensities(1.0, 6000.0)
    want = np.array([0.1124012036712877, 0.13036921243761494, 0.1329718775370217])
    assert np.max(np.abs(got - want)) <= 3e-15
    got = oracle.blackbody_band_intensities(1.0e38, 220.0)
    want = np.array([345.1972853182831, 9.428001052700576e-05, 5.526972080378235e-14])
    assert np.all(np.abs(got - want) <= 3e-14 * want)


def test_oracle_kernel_integral_closes_stefan_boltzmann(oracle):
    assert abs(oracle.planck_kernel_integral(0.0, 80.0) - math.pi ** 4 / 15.0) <= 4e-14


def test_oracle_band_fractions_match_high_precision_quadrature(oracle):
    """Band fractions from 30-digit adaptive quadrature of t^3 / (e^t - 1)
    (mpmath), across the table's temperature range."""
    mpmath.mp.dps = 30
    c2 = mpmath.mpf("0.014387768775039337")
    total = mpmath.pi ** 4 / 15
    bands = [(600, 700), (500, 600), (400, 500)]
    for t in (3.0, 300.0, 2000.0, 6000.0, 4.5e4, 1.0e6, 1.6e7):
        got = oracle.band_log2_fractions(t)
        for b, (lo, hi) in enumerate(bands):
            x0 = c2 / (mpmath.mpf(hi) * mpmath.mpf("1e-9") * t)
            x1 = c2 / (mpmath.mpf(lo) * mpmath.mpf("1e-9") * t)
            # break points near the lower end, where a Wien tail decays over unit width
            pts = [x0] + [x0 + d for d in (0.25, 0.5, 1, 2, 4, 8, 16, 32, 64, 128, 256) if x0 + d < x1]
            val = mpmath.quad(lambda u: u ** 3 / mpmath.expm1(u), pts + [x1])
            want = float(mpmath.log(val / total, 2))
            assert abs(got[b] - want) <= 1e-12 * max(1.0, abs(want)), (t, b, got[b], want)


def test_oracle_fraction_bounds_and_slab_partition(oracle):
    """Fractions in [0, 1] summing to at most 1 over the table's range; a
    constant-source slab split in two transports like the whole
    (radiation.rs:289-327)."""
    rng = np.random.default_rng(5)
    for lt in np.concatenate([[-8.0, 24.0], rng.uniform(-8, 24, 40)]):
        f = np.exp2(oracle.band_log2_fractions(2.0 ** lt))
        assert np.all((f >= 0) & (f <= 1)) and f.sum() <= 1 + 16 * np.finfo(float).eps
    for _ in range(40):
        inc, src = rng.uniform(0, 1e6), rng.uniform(0, 1e6)
        d1, d2 = rng.uniform(0, 20), rng.uniform(0, 20)
        a = oracle.transport_bolometric(inc, oracle.slab(d1, oracle.constant_source_emission(d1, src)))
        b = oracle.transport_bolometric(a[0], oracle.slab(d2, oracle.constant_source_emission(d2, src)))
        c = oracle.transport_bolometric(inc, oracle.slab(d1 + d2,
                                                         oracle.constant_source_emission(d1 + d2, src)))
        assert abs(b[0] - c[0]) <= 32 * np.finfo(float).eps * max(abs(b[0]), abs(c[0]), 1.0)
        assert c[2] <= a[2] and c[2] <= b[2]


# ---------------------------------------------------------------- product host side

def _lib():
    from gravlume_amd import _capi
    return _capi, _capi.load()


def test_product_band_fractions_match_the_oracle(oracle):
    """The product's series evaluation (gl_blackbody_log2_fractions) against
    the oracle's adaptive Simpson, over the table's range and beyond."""
    _capi, L = _lib()
    out = np.zeros(3)
    for lt in np.linspace(-8.0, 24.0, 97):
        t = 2.0 ** lt
        assert L.gl_blackbody_log2_fractions(t, out.ctypes.data_as(C.POINTER(C.c_double))) == 0
        want = oracle.band_log2_fractions(t)
        assert np.max(np.abs(out - want) / np.maximum(1.0, np.abs(want))) <= 1e-12, t
    assert L.gl_blackbody_log2_fractions(-1.0, out.ctypes.data_as(C.POINTER(C.c_double))) != 0
    assert L.gl_blackbody_log2_fractions(float("nan"), out.ctypes.data_as(C.POINTER(C.c_double))) != 0


def test_product_band_table_entries_and_interpolation(oracle):
    """Entries are the f32 roundings of the oracle's log2 fractions at
    T = 2^(-8 + i/128); linear interpolation in log2 T (as the kernels do)
    stays inside the reference's table budgets at every midpoint."""
    _capi, L = _lib()
    n = _capi.BLACKBODY_LUT_ENTRIES
    lut = np.zeros(4 * n, dtype=np.float32)
    assert L.gl_blackbody_lut(lut.ctypes.data_as(C.POINTER(C.c_float)), lut.size) == 0
    assert L.gl_blackbody_lut(lut.ctypes.data_as(C.POINTER(C.c_float)), lut.size - 1) != 0
    lut = lut.reshape(n, 4)
    assert np.all(lut[:, 3] == 0) and np.all(lut[:, :3] <= 0) and np.all(np.isfinite(lut))
    for i in range(0, n, 16):
        want = oracle.band_log2_fractions(2.0 ** (-8 + i / 128))
        assert np.all(lut[i, :3] == want.astype(np.float32)), i
    worst_abs = worst_rel = 0.0
    for i in range(0, n - 1, 3):
        c = i + 0.5
        interp = 0.5 * (lut[i, :3].astype(np.float64) + lut[i + 1, :3].astype(np.float64))
        exact = np.exp2(oracle.band_log2_fractions(2.0 ** (-8 + c / 128)))
        got = np.exp2(interp)
        worst_abs = max(worst_abs, np.max(np.abs(got - exact)))
        vis = exact >= LUT_REL_FLOOR
        if vis.any():
            worst_rel = max(worst_rel, np.max(np.abs(got[vis] / exact[vis] - 1)))
    print("table interpolation: abs", worst_abs, "rel", worst_rel)
    assert worst_abs <= LUT_ABS_FRACTION
    assert worst_rel <= LUT_VISIBLE_REL


def _surface_scene(**kw):
    from gravlume_amd import scene as S
    base = dict(spin=0.8, inclination_deg=60.0, observer_radius=30.0, width=64, height=36,
                fov_deg=45.0, chart=S.INGOING, disk=(6.0, 20.0), escape_radius=200.0,
                escape_map=False, emission="blackbody", temperature_at_six=6000.0)
    base.update(kw)
    return S.Scene(**base)


def _validate(sc):
    _capi, L = _lib()
    d = sc.to_desc()
    st = L.gl_scene_validate(C.byref(d))
    return st, L.gl_last_error().decode()


def test_scene_checks_of_emitter_and_slab():
    """emission.rs / radiation.rs / scene.rs:58-77 / input.rs:166-320 at the
    C ABI: every v3 fixture transport is accepted; a blackbody emitter behind
    a neutral emitting slab, temperatures outside the band table, negative
    depth or emission and an underflowing transmittance are rejected."""
    from gravlume_amd import scene as S
    from gravlume_amd import _capi
    T = S.Transport
    for tr in (T.vacuum(), T.pure_absorption(0.75), T.constant_blackbody(0.35, 0.05, 4500.0),
               T.pure_emission_blackbody(0.003, 4500.0)):
        assert _validate(_surface_scene(transport=tr))[0] == _capi.GL_OK
    assert _validate(_surface_scene(emission="bolometric", temperature_at_six=None,
                                    transport=T.constant_bolometric(0.3, 0.1)))[0] == _capi.GL_OK
    st, msg = _validate(_surface_scene(transport=T.constant_bolometric(0.3, 0.1)))
    assert st == _capi.GL_ERR_DOMAIN and "spectrum" in msg
    st, msg = _validate(_surface_scene(temperature_at_six=3.0e7))
    assert st == _capi.GL_ERR_DOMAIN and "band table" in msg
    st, _ = _validate(_surface_scene(transport=T.pure_emission_blackbody(0.003, 2.0e8)))
    assert st == _capi.GL_ERR_DOMAIN
    st, _ = _validate(_surface_scene(transport=T(True, -0.1, 0.0, None)))
    assert st == _capi.GL_ERR_INVALID_ARGUMENT
    st, _ = _validate(_surface_scene(transport=T(True, 0.1, -1.0, None)))
    assert st == _capi.GL_ERR_INVALID_ARGUMENT
    st, _ = _validate(_surface_scene(transport=T.pure_absorption(120.0)))
    assert st == _capi.GL_ERR_NOT_REPRESENTABLE
    st, _ = _validate(_surface_scene(intensity_at_six=-1.0))
    assert st == _capi.GL_ERR_INVALID_ARGUMENT
