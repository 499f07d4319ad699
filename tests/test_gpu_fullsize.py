"""The benchmark workload at full size -- the 64^3 tet grid (274,625 particles,
1,310,720 tetrahedra) and the 512^2 cloth on a sphere -- through the C ABI.

* The first linearisation (pA, pb, fi) of both bodies is bit-identical to the
  oracle's: same element math, same per-block summation order, same row order.
* In sequential reduction mode two steps of each body reproduce the oracle's
  positions and velocities bit for bit.
* In the default FP64 mode the three forms that bench.py can time -- two
  handles, one batch of the two, captured steps -- give bit-identical
  trajectories, and static vertices do not move.
"""

import numpy as np
import pytest

from bounce_softbody_amd import _lib, configs
from bounce_softbody_amd.body import Body
import oracle

pytestmark = pytest.mark.gpu
SUB, SUB_TOL = 50, 1e-5   # the benchmark's PCG cap and tolerance


@pytest.fixture(scope="module", autouse=True)
def _device():
    if _lib.device_count() < 1:
        pytest.fail("no GPU visible to libb3g")


@pytest.fixture(scope="module")
def bodies():
    return {"tet64": configs.tet_grid(64).arrays(), "cloth511": configs.cloth_on_sphere(511).arrays()}


@pytest.mark.parametrize("name", ["tet64", "cloth511"])
def test_full_size_first_linearisation_bit_exact(bodies, name):
    arr = bodies[name]
    g = Body(arr)
    grow, gcol, gv, gpb, gfi = g.assemble(configs.DT)
    o = oracle.OracleBody(arr)
    orow, ocol, ov, opb, ofi = o.assemble(configs.DT)
    assert np.array_equal(grow, orow) and np.array_equal(gcol, ocol)
    assert np.array_equal(gv, ov) and np.array_equal(gpb, opb) and np.array_equal(gfi, ofi)


@pytest.mark.parametrize("name", ["tet64", "cloth511"])
def test_full_size_sequential_steps_bit_exact(bodies, name):
    arr = bodies[name]
    g = Body(arr)
    g.set_reduction(_lib.REDUCE_SEQUENTIAL)
    o = oracle.OracleBody(arr)
    for step in range(2):
        g.Step(configs.DT, 1, SUB, 0.0, SUB_TOL)
        o.step(configs.DT, 1, SUB, 0.0, SUB_TOL)
        xg, vg = g.state()
        xo, vo = o.state()
        assert np.array_equal(xg, xo) and np.array_equal(vg, vo), f"{name}: step {step}"


def test_full_size_forms_agree(bodies):
    arrs = [bodies["tet64"], bodies["cloth511"]]
    handles = [Body(a) for a in arrs]
    captured = [Body(a) for a in arrs]
    for b in captured:
        b.set_graph(True)
    batch = Body(arrs)
    x0 = [h.state()[0] for h in handles]
    for step in range(3):
        for b in handles + captured + [batch]:
            b.Step(configs.DT, 1, SUB, 0.0, SUB_TOL)
        xb, vb = batch.state()
        for k, (xs, vs) in enumerate(zip(batch.split(xb), batch.split(vb))):
            xh, vh = handles[k].state()
            xc, vc = captured[k].state()
            assert np.array_equal(xs, xh) and np.array_equal(vs, vh), f"batch instance {k}, step {step}"
            assert np.array_equal(xc, xh) and np.array_equal(vc, vh), f"captured body {k}, step {step}"
    for k, a in enumerate(arrs):
        x, v = handles[k].state()
        static = a["p_type"] == 0
        assert static.any()
        assert np.array_equal(x[static], x0[k][static]) and not np.any(v[static])
        assert np.all(np.isfinite(x)) and np.all(np.isfinite(v))
        assert not np.array_equal(x, x0[k])   # the dynamic vertices did move


def test_full_size_precision_forms_reach_fp64():
    """The precision workload of bench.py at full size: a 64^3 tet grid whose Young's moduli are log-uniform in
    [1e4, 1e7] Pa, one step from a deformed, moving state. Both refined forms (FP64 refinement around 50 inner
    iterations at 1e-3, at most 4 sweeps, to an FP64 residual of 1e-4) come within 1e-5 of the FP64 oracle's
    velocities (relative max-norm); the plain FP32 PCG (50 iterations at 1e-5) does not."""
    arr = configs.tet_grid(64, young_per_element=configs.log_uniform_young(64, seed=7)).arrays()
    g = Body(arr)
    for _ in range(10):
        g.Step(configs.DT, 1, 200, 0.0, 1e-6)
    start = g.save_state()
    x, v = g.state()
    o = oracle.OracleBody(arr, precision=64)
    o.set_state(x, v)
    o.step(configs.DT, 1, 20000, 0.0, 1e-10)
    v64 = o.state64()[1].reshape(-1, 3)
    scale = float(np.abs(v64).max())
    errs = {}
    for name, mode, inner, sub, tol in (("plain", _lib.SOLVER_PCG, (), SUB, SUB_TOL),
                                        ("refined_fp32", _lib.SOLVER_REFINED_FP32, (50, 1e-3), 200, 1e-4),
                                        ("refined_bf16", _lib.SOLVER_REFINED_BF16, (50, 1e-3), 200, 1e-4)):
        g.set_solver(mode, *inner)
        g.load_state(start)
        g.Step(configs.DT, 1, sub, 0.0, tol)
        errs[name] = float(np.abs(g.state()[1] - v64).max()) / scale
    print(errs)
    assert errs["refined_fp32"] <= 1e-5 and errs["refined_bf16"] <= 1e-5, errs
    assert errs["plain"] > 1e-5, errs


def test_two_pass_partial_sums_equal_the_separate_form():
    """A body with more than 4096 SpMV tiles (72^3 cells: 4.96 M blocks, at most 1024 per tile; still summed in
    the consuming kernels by default): the vector update's fixed-order sum of the SpMV's partials takes a second
    pass of 16 x 256 after the first one, which it requests before its rows. The fused form equals the
    separate-kernel form bit for bit, step after step."""
    arr = configs.tet_grid(72).arrays()
    sep, fused = Body(arr), Body(arr)
    sep.set_reduction(_lib.REDUCE_FP64_SEPARATE)
    fused.set_reduction(_lib.REDUCE_FP64_FUSED)
    for step in range(3):
        sep.Step(configs.DT, 1, SUB, 0.0, SUB_TOL)
        fused.Step(configs.DT, 1, SUB, 0.0, SUB_TOL)
        xa, va = sep.state()
        xb, vb = fused.state()
        assert np.array_equal(xa, xb) and np.array_equal(va, vb), step
    assert fused.system_size()[1] > 4096 * 1024
    assert [s[0].tolist() for s in sep.instance_stats()] == [s[0].tolist() for s in fused.instance_stats()]
