"""The tetrahedron polar decomposition skips the double sincos for half-angles below 2e-4 and uses theta and 1:
tests/native/small_angle_check.cpp checks every float in [0, 2e-4) against the host's libm (the one the reference's
build calls), (float)sin((double)t) == t and (float)cos((double)t) == 1."""

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_small_angle_sin_cos_narrow_exactly(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    exe = str(tmp_path / "small_angle_check")
    subprocess.run([gxx, "-O2", "-std=c++17", "-pthread", os.path.join(ROOT, "tests", "native", "small_angle_check.cpp"),
                    "-o", exe], check=True, timeout=300)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "counterexamples 0" in r.stdout
