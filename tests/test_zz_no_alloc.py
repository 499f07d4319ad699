"""b3g_solve_cg_static: the PCG on device arrays with its scratch in a fixed device array of the library (no
allocation), against b3g_solve_cg (host arrays, scratch allocated per call) on the same systems.

Both run the same kernels; they cut the rows into different tiles, so their FP64 dot-product partials group the
same float terms differently. The solutions agree to the floating-point bar of the drop-in (checked here at 1e-5
relative max-norm, iteration counts within one). Two calls in a row, the second on a smaller system, show that
nothing of the first call's scratch contents leaks into the second; a malformed matrix (a column index past the
last row, planted on the host) is refused on the device before any PCG kernel reads it.
"""

import numpy as np
import pytest

from bounce_softbody_amd import _lib, configs
from bounce_softbody_amd.body import Body

pytestmark = pytest.mark.gpu
MAX_IT, TOL = 200, 1e-5


@pytest.fixture(scope="module", autouse=True)
def _device():
    if _lib.device_count() < 1:
        pytest.fail("no GPU visible to libb3g")


def system(N):
    """pA and pb of the first linearisation of an N^3 tet grid (list-order CSR, column-major blocks)."""
    rowptr, cols, vals, pb, _ = Body(configs.tet_grid(N).arrays()).assemble(configs.DT)
    return rowptr, cols, vals, pb


def static_solve(rowptr, cols, vals, pb, x0):
    n, nnz = len(rowptr) - 1, len(cols)
    planes = np.ascontiguousarray(np.asarray(vals, np.float32).reshape(nnz, 9).T)   # entry c of block k at c * nnz + k
    d_rowptr, d_cols = _lib.DeviceArray.of(np.asarray(rowptr, np.uint32)), _lib.DeviceArray.of(np.asarray(cols, np.uint32))
    d_vals, d_b = _lib.DeviceArray.of(planes), _lib.DeviceArray.of(np.asarray(pb, np.float32))
    d_x, d_it = _lib.DeviceArray.of(np.asarray(x0, np.float32)), _lib.DeviceArray.of(np.zeros(1, np.uint32))
    _lib.solve_cg_static(n, nnz, d_rowptr, d_cols, d_vals, d_b, d_x, MAX_IT, TOL, d_it)
    return d_x.download(np.float32, (n, 3)), int(d_it.download(np.uint32, (1,))[0])


def rel(a, b):
    return float(np.abs(a - b).max() / max(float(np.abs(b).max()), 1e-30))


def test_static_solve_equals_allocating_solve_twice():
    for N, seed in ((14, 1), (6, 2)):   # 3375 rows, then 343
        rowptr, cols, vals, pb = system(N)
        x0 = np.random.default_rng(seed).normal(scale=1e-3, size=pb.shape).astype(np.float32)
        x_ref, it_ref = _lib.solve_cg(rowptr, cols, vals, pb, x0, MAX_IT, TOL)
        x, it = static_solve(rowptr, cols, vals, pb, x0)
        err = rel(x, x_ref)
        print(f"N={N}: rows {len(rowptr) - 1}, iterations {it} (allocating entry {it_ref}), rel diff {err:.2e}, "
              f"bit-exact {np.array_equal(x, x_ref)}")
        assert abs(it - it_ref) <= 1 and 0 < it <= MAX_IT
        assert err <= 1e-5


def test_static_solve_rejects_malformed_matrix_on_device():
    rowptr, cols, vals, pb = system(4)
    bad = np.array(cols, np.uint32)
    bad[len(bad) // 2] = len(rowptr) + 5     # a column past the last row
    x0 = np.full(pb.shape, 0.25, np.float32)
    x, it = static_solve(rowptr, bad, vals, pb, x0)
    assert it == _lib.CG_STATIC_INVALID
    assert np.array_equal(x, x0)             # left unchanged
    x, it = static_solve(rowptr, cols, vals, pb, np.zeros_like(pb))   # the next call is unaffected
    x_ref, it_ref = _lib.solve_cg(rowptr, cols, vals, pb, np.zeros_like(pb), MAX_IT, TOL)
    assert abs(it - it_ref) <= 1 and rel(x, x_ref) <= 1e-5


def test_static_solve_limits_are_host_errors():
    one = _lib.DeviceArray.of(np.zeros(4, np.uint32))
    with pytest.raises(_lib.B3gError, match="rows"):
        _lib.solve_cg_static(_lib.CG_STATIC_MAX_ROWS + 1, 1, one, one, one, one, one, 10, TOL)
    with pytest.raises(_lib.B3gError, match="max_iterations"):
        _lib.solve_cg_static(1, 1, one, one, one, one, one, _lib.CG_STATIC_MAX_ITERATIONS + 1, TOL)
    host = np.zeros(16, np.float32)
    with pytest.raises(_lib.B3gError, match="not device memory"):
        _lib.solve_cg_static(1, 1, one, one, host.ctypes.data, one, one, 10, TOL)
