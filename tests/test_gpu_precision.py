"""Precision form: FP64 iterative refinement around an FP32 / BF16 inner PCG.

Errors are stated against the FP64 oracle (the reference restated with its
scalar type set to double, pinned bit for bit to the reference compiled with
scalar = double by tests/test_oracle_golden.py, including this file's FP64
step of the stiffness-contrast grid) stepping from the identical state with a
tight PCG tolerance: relative max-norm error of the step's velocities. The scene
is the stiffness-contrast workload at small size: a tet grid whose elements
have Young's moduli log-uniform in [1e4, 1e7] Pa.
"""

import numpy as np
import pytest

from bounce_softbody_amd import _lib, configs
from bounce_softbody_amd.body import Body
import oracle

pytestmark = pytest.mark.gpu

N = 8
TARGET = 1e-5          # relative error vs the FP64 oracle the refined forms must reach
FP64_TOL = 1e-10       # oracle PCG tolerance (FP64)


@pytest.fixture(scope="module", autouse=True)
def _device():
    if _lib.device_count() < 1:
        pytest.fail("no GPU visible to libb3g")


def scene():
    return configs.tet_grid(N, young_per_element=configs.log_uniform_young(N, seed=7)).arrays()


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def fp64_step(arr, x, v):
    o = oracle.OracleBody(arr, precision=64)
    o.set_state(x, v)
    o.step(configs.DT, 1, 20000, 0.0, FP64_TOL)
    return o.state64()


def gpu_step(arr, x, v, mode, sub_iterations, sub_tolerance, inner=(20, 1e-2)):
    g = Body(arr)
    g.set_solver(mode, *inner)
    g.set_state(x, v)
    g.Step(configs.DT, 1, sub_iterations, 0.0, sub_tolerance)
    return g.state(), g


def states():
    """The rest state and a deformed, moving state (after 15 plain steps)."""
    arr = scene()
    g = Body(arr)
    x0, v0 = g.state()
    for _ in range(15):
        g.Step(configs.DT, 1, 200, 0.0, 1e-6)
    x1, v1 = g.state()
    return arr, [(x0, v0), (x1, v1)]


@pytest.mark.parametrize("mode", [_lib.SOLVER_REFINED_FP32, _lib.SOLVER_REFINED_BF16], ids=["refined_fp32", "refined_bf16"])
def test_refined_forms_reach_fp64_oracle(mode):
    arr, sts = states()
    for x, v in sts:
        _, v64 = fp64_step(arr, x, v)
        (_, vg), g = gpu_step(arr, x, v, mode, 600, 1e-9)
        err = rel(vg, v64)
        sweeps, res = g.refinement()[0]
        print(f"mode {mode}: rel err vs FP64 {err:.2e}, sweeps {sweeps}, FP64 residual {res:.2e}")
        assert sweeps >= 1 and res <= 1e-6
        assert err <= TARGET, f"{err:.2e}"


def test_plain_fp32_error_reported():
    """The reference PCG in FP32 at tolerance 1e-5: its error vs FP64 (reported, bounded loosely)."""
    arr, sts = states()
    for x, v in sts:
        _, v64 = fp64_step(arr, x, v)
        (_, vg), g = gpu_step(arr, x, v, _lib.SOLVER_PCG, 2000, 1e-5)
        err = rel(vg, v64)
        print(f"plain FP32 PCG tol 1e-5: rel err vs FP64 {err:.2e}, iterations {g.stats()[0][2]}")
        assert err < 1e-2


def test_refined_mode_keeps_plain_semantics_when_off():
    """Switching back to the reference PCG gives the plain path's results bit for bit."""
    arr = scene()
    a, b = Body(arr), Body(arr)
    b.set_solver(_lib.SOLVER_REFINED_BF16, 10, 1e-2)
    b.Step(configs.DT, 1, 100, 0.0, 1e-8)
    b.set_solver(_lib.SOLVER_PCG)
    b.set_state(*a.state())
    b.load_state(a.save_state())
    for _ in range(3):
        a.Step(configs.DT, 1, 50, 0.0, 1e-5)
        b.Step(configs.DT, 1, 50, 0.0, 1e-5)
    xa, va = a.state()
    xb, vb = b.state()
    assert np.array_equal(xa, xb) and np.array_equal(va, vb)


def test_refined_batch_equals_standalone():
    arrs = [scene(), configs.PARITY_SCENES["cloth_sphere"]().arrays(), configs.tet_grid(4).arrays()]
    singles = [Body(a) for a in arrs]
    batch = Body(arrs)
    for body in singles + [batch]:
        body.set_solver(_lib.SOLVER_REFINED_BF16, 15, 1e-2)
    for _ in range(4):
        for s in singles:
            s.Step(configs.DT, 1, 300, 0.0, 1e-8)
        batch.Step(configs.DT, 1, 300, 0.0, 1e-8)
    xb, vb = batch.state()
    for k, (xs, vs) in enumerate(zip(batch.split(xb), batch.split(vb))):
        x1, v1 = singles[k].state()
        assert np.array_equal(xs, x1) and np.array_equal(vs, v1), f"instance {k}"


def test_set_solver_validates():
    b = Body(configs.tet_grid(2).arrays())
    with pytest.raises(_lib.B3gError):
        b.set_solver(7)
    with pytest.raises(_lib.B3gError):
        b.set_solver(_lib.SOLVER_REFINED_FP32, 0, 1e-2)
    with pytest.raises(_lib.B3gError):
        b.set_solver(_lib.SOLVER_REFINED_FP32, 10, 1.5)
