"""Parity of the configuration bench.py times, at its full size, against the oracle.

bench.py steps the 64^3 tet grid (274,625 particles, 1,310,720 tetrahedra) and the
512^2 cloth on a sphere with one Newton iteration and at most 50 PCG iterations at
tolerance 1e-5, in the default FP64 reduction mode. Per step, for 10 consecutive
steps of each body:

* a sequential-mode body (the reference's FP32 summation order) reproduces the
  oracle's positions, velocities and solver statistics bit for bit, so its state is
  the oracle's state at every step;
* from that same state (checkpoint), the default-mode body's step is within 1e-4
  relative max-norm of the oracle's step restated with the default mode's one
  documented difference -- the PCG dot products and the Newton error accumulated in
  double (``OracleBody.set_wide_dots``) -- with the same Newton count and PCG counts
  within one of the reference's;
* against the reference's own FP32 accumulation the step is within 1e-4 whenever
  the PCG converged before its cap. A capped, unconverged PCG (the cloth once it
  rests on the sphere) amplifies the rounding of the accumulator: there the
  reference's own step moves by as much when only its accumulator is widened, and
  the device's deviation from the reference stays within that amount + 1e-4.

The same holds for the first, middle and last bodies of the 4096-body batch
(sparse_solver.cpp:28-109, force_solver.cpp:88-180 for the solve being compared).
"""

import numpy as np
import pytest

from bounce_softbody_amd import _lib, configs
from bounce_softbody_amd.body import Body
import oracle

pytestmark = pytest.mark.gpu
SUB, SUB_TOL = 50, 1e-5   # the benchmark's PCG cap and tolerance
BOUND = 1e-4              # relative max-norm per step


@pytest.fixture(scope="module", autouse=True)
def _device():
    if _lib.device_count() < 1:
        pytest.fail("no GPU visible to libb3g")


def rel(a, b):
    return float(np.abs(a - b).max() / max(float(np.abs(b).max()), 1e-30))


def oracle_pair_step(o):
    """Steps the oracle `o` and a clone of it whose accumulators are widened. Returns the clone's state, whether
    its PCG reached the cap in this step, and the reference-vs-widened deviation (x, v)."""
    w = o.clone()
    w.set_wide_dots(True)
    w.reset_stats()
    o.step(configs.DT, 1, SUB, 0.0, SUB_TOL)
    w.step(configs.DT, 1, SUB, 0.0, SUB_TOL)
    xw, vw = w.state()
    xo, vo = o.state()
    return (xw, vw), int(w.stats()[2]) >= SUB, (rel(xw, xo), rel(vw, vo))


def check_default_step(name, xd, vd, sd, o, wide, capped, spread):
    xo, vo = o.state()
    xw, vw = wide
    so = o.stats()
    ewx, ewv = rel(xd, xw), rel(vd, vw)
    assert ewx <= BOUND and ewv <= BOUND, f"{name}: vs widened oracle x {ewx:.2e} v {ewv:.2e}"
    ex, ev = rel(xd, xo), rel(vd, vo)
    if capped:
        assert ex <= spread[0] + BOUND and ev <= spread[1] + BOUND, \
            f"{name}: capped PCG, vs oracle x {ex:.2e} v {ev:.2e}, reference spread x {spread[0]:.2e} v {spread[1]:.2e}"
    else:
        assert ex <= BOUND and ev <= BOUND, f"{name}: vs oracle x {ex:.2e} v {ev:.2e}"
    assert int(sd[0]) == int(so[0]), f"{name}: Newton iterations {sd} vs {so}"
    assert abs(int(sd[1]) - int(so[1])) <= 1 and abs(int(sd[2]) - int(so[2])) <= 1, f"{name}: PCG {sd} vs {so}"
    exact = np.array_equal(xd, xw) and np.array_equal(vd, vw)
    return (ex, ev), (ewx, ewv), exact


@pytest.mark.parametrize("name", ["tet64", "cloth511"])
def test_default_mode_steps_vs_oracle_full_size(name):
    arr = configs.tet_grid(64).arrays() if name == "tet64" else configs.cloth_on_sphere(511).arrays()
    seq = Body(arr)
    seq.set_reduction(_lib.REDUCE_SEQUENTIAL)
    dflt = Body(arr)
    o = oracle.OracleBody(arr)
    for step in range(10):
        start = seq.save_state()
        wide, capped, spread = oracle_pair_step(o)
        seq.Step(configs.DT, 1, SUB, 0.0, SUB_TOL)
        xo, vo = o.state()
        xs, vs = seq.state()
        assert np.array_equal(xs, xo) and np.array_equal(vs, vo), f"{name}: sequential step {step}"
        assert np.array_equal(seq.stats()[0], o.stats()), f"{name}: sequential statistics, step {step}"
        dflt.load_state(start)
        dflt.Step(configs.DT, 1, SUB, 0.0, SUB_TOL)
        xd, vd = dflt.state()
        (ex, ev), (ewx, ewv), exact = check_default_step(f"{name} step {step}", xd, vd, dflt.stats()[0], o, wide,
                                                          capped, spread)
        print(f"{name} step {step}: PCG max {o.stats()[2]}{' (capped)' if capped else ''}; default mode vs oracle "
              f"x {ex:.2e} v {ev:.2e}; vs widened oracle x {ewx:.2e} v {ewv:.2e}{' (bit-exact)' if exact else ''}; "
              f"reference vs widened x {spread[0]:.2e} v {spread[1]:.2e}", flush=True)


def test_batch4096_instances_vs_oracle():
    """The 4096-body batch (12.6 M particles, 148.8 M blocks): instances 0, 2047 and 4095 against their oracle
    bodies for 3 steps, bit for bit in sequential mode and, in the default mode, per step as above."""
    arrs = configs.batch_boxes(4096)
    picks = [0, 2047, 4095]
    batch = Body(arrs)
    oracles = [oracle.OracleBody(arrs[k]) for k in picks]
    for step in range(3):
        start = batch.save_state()
        batch.set_reduction(_lib.REDUCE_SEQUENTIAL)
        batch.Step(configs.DT, 1, SUB, 0.0, SUB_TOL)
        pairs = [oracle_pair_step(o) for o in oracles]
        xb, vb = batch.state()
        xs, vs = batch.split(xb), batch.split(vb)
        stats = batch.instance_stats()
        for o, k in zip(oracles, picks):
            xo, vo = o.state()
            assert np.array_equal(xs[k], xo) and np.array_equal(vs[k], vo), f"instance {k}: sequential step {step}"
            assert np.array_equal(stats[k][0], o.stats()), f"instance {k}: sequential statistics, step {step}"
        after = batch.save_state()
        batch.load_state(start)
        batch.set_reduction(_lib.REDUCE_FP64)
        batch.Step(configs.DT, 1, SUB, 0.0, SUB_TOL)
        xb, vb = batch.state()
        xs, vs = batch.split(xb), batch.split(vb)
        stats = batch.instance_stats()
        for o, k, (wide, capped, spread) in zip(oracles, picks, pairs):
            (ex, ev), (ewx, ewv), exact = check_default_step(f"instance {k} step {step}", xs[k], vs[k], stats[k][0],
                                                              o, wide, capped, spread)
            print(f"instance {k} step {step}: PCG max {o.stats()[2]}; default vs oracle x {ex:.2e} v {ev:.2e}; "
                  f"vs widened x {ewx:.2e} v {ewv:.2e}{' (bit-exact)' if exact else ''}", flush=True)
        batch.load_state(after)
