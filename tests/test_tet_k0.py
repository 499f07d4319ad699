"""The device's recomputed tetrahedron stiffness blocks (bounce_softbody_amd/csrc/b3g_tet.h) equal the oracle's
dense K0 = V B^T D B (tetrahedron_element_force.cpp:219-265) bit for bit.

tests/native/tet_k0_check.cpp includes the product header and the oracle and compares all 16 blocks of 200,000
seeded elements (jittered, axis-aligned with zero gradient entries, random, slivers; E over five decades); built
with g++ for the host, no GPU involved.
"""

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_recomputed_k0_equals_dense_reference(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    exe = str(tmp_path / "tet_k0_check")
    cmd = [gxx, "-std=c++17", "-O2", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           "-I", os.path.join(ROOT, "bounce_softbody_amd", "csrc"), "-I", os.path.join(ROOT, "oracle"),
           os.path.join(ROOT, "tests", "native", "tet_k0_check.cpp"), "-o", exe]
    subprocess.run(cmd, check=True, timeout=300)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("mismatches 0")
    assert int(r.stdout.split()[1]) > 20_000_000
