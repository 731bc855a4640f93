"""Host (CPU) check of the equatorial-crossing root on the f32 quartic dense
output (z_root of gravlume_amd/csrc/trace_device.h, compiled with g++):
the binary32 root stays within its binary32 condition number of the binary64
root of the same step (tests/cpp/dense_root_check.cpp)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_f32_crossing_root_matches_the_f64_root(tmp_path):
    exe = tmp_path / "dense_root_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", str(exe),
                    os.path.join(HERE, "cpp", "dense_root_check.cpp")], check=True)
    res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(res.stdout)
    assert res.returncode == 0, res.stdout
