"""The C-ABI library loads, exports every entry point include/gravlume.h
declares, and its host-side validation mirrors the reference's
GpuTraceInputError cases (crates/gravlume-render/src/trace/input.rs:342-381).
No kernel runs here."""
import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gravlume.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:gl_status|const char\*)\s+(gl_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    from gravlume_amd import _capi
    lib_path = _capi.LIB_PATH
    assert os.path.exists(lib_path), "build libgravlume_hip.so first"
    out = subprocess.run(["nm", "-D", "--defined-only", lib_path], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (gl_\w+)$", out, re.M))
    declared = declared_functions()
    assert len(declared) >= 18
    missing = [f for f in declared if f not in exported]
    assert not missing, missing
    # and the Python binding covers exactly the declared set
    assert sorted(_capi.SIGNATURES) == declared


def test_struct_layouts_match_the_header(tmp_path):
    """ctypes mirrors agree with the C compiler's layout of include/gravlume.h."""
    from gravlume_amd import _capi
    structs = {"gl_scene_desc": _capi.SceneDesc, "gl_ray_outputs": _capi.RayOutputsC,
               "gl_tile_region": _capi.TileRegionC, "gl_persistent_config": _capi.PersistentConfig,
               "gl_queue_stats": _capi.QueueStats, "gl_work_item": _capi.WorkItem}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for field, _ in py._fields_:
            lines.append(f'printf("{cname}.{field} %zu\\n", offsetof({cname}, {field}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-o", str(exe), str(src)], check=True)
    got = dict(line.rsplit(" ", 1) for line in subprocess.run(
        [str(exe)], capture_output=True, text=True, check=True).stdout.splitlines())
    for cname, py in structs.items():
        assert int(got[cname]) == C.sizeof(py), cname
        for field, _ in py._fields_:
            assert int(got[f"{cname}.{field}"]) == getattr(py, field).offset, (cname, field)


def test_flop_constants_are_counted_from_the_kernel_source():
    import gravlume_amd
    rhs, attempt = gravlume_amd.flop_constants()
    assert 80 <= rhs <= 200
    assert 6 * rhs < attempt < 6 * rhs + 800


def _plan_error(**changes):
    import gravlume_amd
    from gravlume_amd import scene as S
    sc = S.small_cpu_scene(**changes)
    with pytest.raises(gravlume_amd.GpuTraceInputError) as e:
        gravlume_amd.TracePipeline([sc], escape_map_rgb=None) if not sc.escape_map else \
            gravlume_amd.TracePipeline([sc])
    return str(e.value)


def test_plan_rejects_non_normalized_or_invalid_scenes(monkeypatch):
    import torch
    monkeypatch.setattr(torch.cuda, "init", lambda: None)
    assert "inside the escape radius" in _plan_error(disk=(None, 1000.0))
    assert "field of view" in _plan_error(fov_deg=180.0)
    assert "g_tt" in _plan_error(observer_radius=1.9, inclination_deg=90.0, spin=0.0, charge=0.0)
    assert "step policy" in _plan_error(minimum_step=1.0, maximum_step=0.5)
    assert "surface radii" in _plan_error(disk=(5.0, 4.0))
    assert "without a surface" in _plan_error(capture_certificate=True)


def test_mass_and_frequency_normalization_is_required(monkeypatch):
    import torch
    import gravlume_amd
    from gravlume_amd import _capi, scene as S
    monkeypatch.setattr(torch.cuda, "init", lambda: None)
    lib = _capi.load()
    d = S.small_cpu_scene(escape_map=False).to_desc()
    d.mass = 1.0 + 2.0 ** -52
    h = C.c_void_p()
    st = lib.gl_trace_plan_create(C.byref(d), 1, None, 0, 0, C.byref(h))
    assert st == 2 and b"mass must be exactly 1" in lib.gl_last_error()
    d = S.small_cpu_scene(escape_map=False).to_desc()
    d.frequency = 1.0 + 1e-12
    st = lib.gl_trace_plan_create(C.byref(d), 1, None, 0, 0, C.byref(h))
    assert st == 2 and b"frequency" in lib.gl_last_error()
    d = S.small_cpu_scene(escape_map=True).to_desc()
    st = lib.gl_trace_plan_create(C.byref(d), 1, None, 0, 0, C.byref(h))
    assert st == 1 and b"escape-map" in lib.gl_last_error()
    assert lib.gl_trace_plan_create(None, 0, None, 0, 0, C.byref(h)) == 1


def test_extremality_change_under_f32_packing_is_rejected(monkeypatch):
    import torch
    from gravlume_amd import _capi, scene as S
    monkeypatch.setattr(torch.cuda, "init", lambda: None)
    lib = _capi.load()
    d = S.small_cpu_scene(escape_map=False, disk=None).to_desc()
    d.spin = 1.0 - 2.0 ** -40  # subextremal in f64, extremal once rounded to f32
    d.charge = 0.0
    h = C.c_void_p()
    st = lib.gl_trace_plan_create(C.byref(d), 1, None, 0, 0, C.byref(h))
    assert st == 3 and b"extremality" in lib.gl_last_error()


def test_product_isco_matches_oracle(oracle):
    from gravlume_amd import isco_prograde
    for a, q in [(0.94, 0.0), (0.5, 0.3), (0.9, 0.2), (0.92, 0.18), (0.998, 0.05), (0.0, 1.0),
                 (0.3, 0.9), (0.0, 0.0)]:
        assert abs(isco_prograde(a, q) - oracle.isco_prograde(a, q)) <= 1e-12


def test_scene_validate_needs_no_device_and_matches_the_oracle_frame_check(oracle):
    """gl_scene_validate runs the plan's scene checks on the host, with the
    oracle's verdict on the observer frame: a draw whose right-leg axes tie
    in their max-norm score (an axis nearly inside the span of the sight and
    up legs normalizes to the same direction as its neighbour) builds in both
    within the 1e-12 Gram budget; an observer inside the ergosphere (g_tt >= 0)
    is refused by both."""
    import math
    from gravlume_amd import _capi, scene as S
    lib = _capi.load()
    good = S.small_cpu_scene(escape_map=False)
    assert lib.gl_scene_validate(C.byref(good.to_desc())) == 0
    tie = S.Scene(spin=0.08186844760218794, charge=0.7823507233737966,
                  inclination_deg=74.81317707591471, observer_radius=196.28397290333865,
                  width=512, height=512, fov_deg=20.0, disk=(None, 20.0), escape_map=False)
    assert lib.gl_scene_validate(C.byref(tie.to_desc())) == 0, lib.gl_last_error()
    o = oracle.observation(1.0, tie.spin, tie.charge, float(tie.chart), [0.0, *tie.observer_xyz()],
                           [0, 0, 0, 0], [0, 0, 1], 1.0, tie.width, tie.height,
                           math.radians(tie.fov_deg))
    assert o.gram_residual <= 1e-12
    ergo = S.Scene(spin=0.9, charge=0.0, inclination_deg=90.0, observer_radius=1.9, width=64,
                   height=64, fov_deg=20.0, disk=None, escape_map=False)
    assert lib.gl_scene_validate(C.byref(ergo.to_desc())) == _capi.GL_ERR_DOMAIN
    assert b"g_tt" in lib.gl_last_error()
    with pytest.raises(ValueError):
        oracle.observation(1.0, ergo.spin, ergo.charge, float(ergo.chart),
                           [0.0, *ergo.observer_xyz()], [0, 0, 0, 0], [0, 0, 1], 1.0, 64, 64,
                           math.radians(ergo.fov_deg))


def test_step_controller_factors_are_validated():
    """The reference policy's controller factors (policy.rs:38-66): 0 < safety
    <= 1, 0 < min <= 1 <= max; zero selects the reference defaults."""
    from gravlume_amd import _capi, scene as S
    lib = _capi.load()
    good = S.small_cpu_scene(escape_map=False)
    d = good.to_desc()
    d.step_safety = d.step_min_factor = d.step_max_factor = 0.0
    assert lib.gl_scene_validate(C.byref(d)) == 0
    for field, value in (("step_safety", 1.5), ("step_min_factor", -0.1), ("step_max_factor", 0.5),
                         ("step_safety", float("nan"))):
        d = good.to_desc()
        setattr(d, field, value)
        assert lib.gl_scene_validate(C.byref(d)) == _capi.GL_ERR_INVALID_ARGUMENT, field


def test_viewport_limits():
    """Sides up to 65536 and pixel counts up to 2^32 - 1 are accepted (32-bit
    pixel indices); 65536 x 65536 = 2^32 pixels is a capacity error, as the
    reference's checked width x height product (scientific_capture.rs:294)."""
    from gravlume_amd import _capi, scene as S
    lib = _capi.load()
    for (w, h), want in (((65536, 65535), 0), ((65535, 65536), 0), ((1, 65536), 0),
                         ((65536, 65536), _capi.GL_ERR_CAPACITY), ((65537, 1), _capi.GL_ERR_CAPACITY),
                         ((0, 8), _capi.GL_ERR_INVALID_ARGUMENT)):
        d = S.small_cpu_scene(escape_map=False, width=w, height=h).to_desc()
        assert lib.gl_scene_validate(C.byref(d)) == want, (w, h, lib.gl_last_error())
