"""The device SDF builder's tree walk (k_sdf_tree) gives every voxel the value of the brute-force loop over all
triangles, bit for bit: tests/native/sdf_tree_check.cpp compiles the builder's host/device functions for the host
(hipcc) and compares both on four meshes (sphere, torus, height field, and a sphere scaled to 30 m, whose far
points lie ~3e6 m away). The brute-force loop itself is pinned to the reference by tests/test_gpu_sdf.py."""

import os
import shutil
import subprocess

import numpy as np
import pytest

from bounce_softbody_amd import configs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _write(path, v, t):
    v = np.ascontiguousarray(v, np.float32).reshape(-1, 3)
    t = np.asarray(t, np.uint32)
    t = np.ascontiguousarray(t.reshape(len(t), -1))
    with open(path, "wb") as f:
        np.array([len(v), len(t), t.shape[1]], np.uint32).tofile(f)
        v.tofile(f)
        t.tofile(f)


def test_tree_walk_equals_brute_force(tmp_path):
    hipcc = shutil.which("hipcc") or ("/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else None)
    if hipcc is None:
        pytest.skip("hipcc not found")
    meshes = {"sphere": configs.icosphere(3, 0.5), "torus": configs.torus_mesh(0.6, 0.2, 24, 12),
              "terrain": configs.terrain_mesh(16, 5, size=2.0, amp=0.3), "big_sphere": configs.icosphere(2, 30.0)}
    files = []
    for name, (v, t) in meshes.items():
        files.append(str(tmp_path / f"{name}.bin"))
        _write(files[-1], v, t)
    exe = str(tmp_path / "sdf_tree_check")
    csrc = os.path.join(ROOT, "bounce_softbody_amd", "csrc")
    cmd = [hipcc, "-x", "hip", "--offload-arch=gfx950", "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
           "-fhip-fp32-correctly-rounded-divide-sqrt", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-I", os.path.join(ROOT, "include"), "-I", csrc, os.path.join(ROOT, "tests", "native", "sdf_tree_check.cpp"),
           "-o", exe]
    subprocess.run(cmd, check=True, timeout=600, capture_output=True)
    r = subprocess.run([exe, "24", *files], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "total mismatches 0" in r.stdout
