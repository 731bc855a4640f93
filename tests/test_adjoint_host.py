"""Host (CPU) checks of the discrete adjoint templates of gravlume_amd/csrc/adjoint.h,
compiled with g++ in double: the reverse-mode RHS product and the frozen-step
trajectory adjoint through an equator crossing, both against central
differences of the same forward maps (tests/cpp/adjoint_check.cpp)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_adjoint_templates_match_finite_differences(tmp_path):
    exe = tmp_path / "adjoint_check"
    subprocess.run(["g++", "-O1", "-std=c++17", "-o", str(exe),
                    os.path.join(HERE, "cpp", "adjoint_check.cpp")], check=True)
    res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(res.stdout)
    assert res.returncode == 0, res.stdout
    assert "rhs_vjp cases" in res.stdout and "trajectory worst" in res.stdout
