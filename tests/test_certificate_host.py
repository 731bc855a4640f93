"""Interval capture certificate (gravlume_amd/csrc/certificate.h) on the host,
against the f64 oracle (tests/cpp/certificate_check.cpp, compiled with the
oracle sources):

* soundness: no certified ray of four scenes (Kerr a = 0.8 default view,
  Kerr-Newman a = 0.5 Q = 0.3 near field, Schwarzschild, Kerr a = 0.9) ends
  anywhere but the horizon in the oracle's strict-v1 trace;
* acceptance: on the reference's default Kerr view (a = 0.8, r = 30 M,
  60 deg, FOV 45 deg, 1280x720, every 4th pixel each way) the certified
  fraction matches the reference's 56,937 / 921,600 px = 6.178%
  (docs/research/gpu-geodesic-acceleration.md:200) to within 0.1 point."""
import os
import re
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_certificate_is_sound_and_matches_reference_acceptance(tmp_path):
    exe = tmp_path / "certificate_check"
    srcs = [os.path.join(HERE, "cpp", "certificate_check.cpp")] + [
        os.path.join(ROOT, "oracle", f) for f in ("spacetime.cpp", "observer.cpp", "trace.cpp",
                                                  "tangent.cpp")]
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", ROOT, "-o", str(exe), *srcs, "-lpthread"],
                   check=True)
    res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    print(res.stdout)
    assert res.returncode == 0, res.stdout
    m = re.search(r"cert default-kerr sampled (\d+) certified (\d+)", res.stdout)
    frac = int(m.group(2)) / int(m.group(1))
    assert abs(frac - 56937 / 921600) <= 1e-3, frac
