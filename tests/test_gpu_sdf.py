"""b3BuildSDF on the device (b3g_build_sdf) against the reference's goldens and the oracle: bit for bit."""

import os

import numpy as np
import pytest

from bounce_softbody_amd import _lib, configs
import oracle
from tests.golden.make_golden import sdf_cases

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "sdf_build.npz")


@pytest.fixture(scope="module", autouse=True)
def _device():
    if _lib.device_count() < 1:
        pytest.fail("no GPU visible to libb3g")


@pytest.mark.parametrize("name", list(sdf_cases()))
def test_sdf_build_matches_reference_golden(name):
    g = np.load(GOLDEN)
    lo, hi, dims, vox = _lib.build_sdf(g[f"{name}/vertices"], g[f"{name}/triangles"], g[f"{name}/cell"],
                                       float(g[f"{name}/extension"]))
    assert np.array_equal(lo, g[f"{name}/lower"]) and np.array_equal(hi, g[f"{name}/upper"])
    assert np.array_equal(vox, g[f"{name}/voxels"])


def test_sdf_build_larger_grid_equals_oracle():
    """A 320-triangle icosphere on a 71^3 grid (357,911 voxels), triangles given as b3Triangle (stride 6)."""
    v, t = configs.icosphere(2, 0.5)
    t6 = np.concatenate([t, np.full(t.shape, 0xFFFFFFFF, np.uint32)], 1)
    a = _lib.build_sdf(v, t6, (0.02, 0.02, 0.02), 0.2)
    b = oracle.build_sdf(v, t, (0.02, 0.02, 0.02), 0.2)
    assert a[2] == b[2] and a[2][0] >= 70
    assert np.array_equal(a[3], b[3])


def test_sdf_build_validates():
    v, t = configs.icosphere(0, 0.5)
    c = np.float32([0.1, 0.1, 0.1])
    out = np.zeros(10, np.float32)
    L = _lib.lib()
    f32p, u32p = _lib.f32p, _lib.u32p
    tt = np.ascontiguousarray(t, np.uint32).reshape(-1)
    vv = np.ascontiguousarray(v, np.float32).reshape(-1)
    st = L.b3g_build_sdf(0, _lib._ptr(vv, f32p), len(v), _lib._ptr(tt, u32p), len(t), 3, _lib._ptr(c, f32p), 0.2,
                         _lib._ptr(out, f32p), out.size)
    assert st == 1 and "grid has" in L.b3g_last_error().decode()   # capacity below W * H * D
    bad = tt.copy()
    bad[4] = 1000
    big = np.zeros(100000, np.float32)
    st = L.b3g_build_sdf(0, _lib._ptr(vv, f32p), len(v), _lib._ptr(bad, u32p), len(t), 3, _lib._ptr(c, f32p), 0.2,
                         _lib._ptr(big, f32p), big.size)
    assert st == 1 and "out of range" in L.b3g_last_error().decode()
