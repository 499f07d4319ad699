"""Captured steps: a HIP graph of the step's launch sequence, replayed, gives
the same trajectories as the uncaptured step, bit for bit. Every replay is
checked, with host reads and writes of the body between replays."""

import numpy as np
import pytest

from bounce_softbody_amd import _lib, configs
from bounce_softbody_amd.body import Body
from tests.golden.make_golden import external_loads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    if _lib.device_count() < 1:
        pytest.fail("no GPU visible to libb3g")


def pair(arr, **solver):
    a, b = Body(arr), Body(arr)
    for body in (a, b):
        if solver:
            body.set_solver(**solver)
    b.set_graph(True)
    return a, b


def check_steps(a, b, steps, sub=30, tol=1e-5, ext=None):
    for k in range(steps):
        if ext is not None and k % 3 == 0:
            a.apply_external(*ext)
            b.apply_external(*ext)
        a.Step(configs.DT, 2, sub, 0.0, tol)
        b.Step(configs.DT, 2, sub, 0.0, tol)
        xa, va = a.state()
        xb, vb = b.state()
        assert np.array_equal(xa, xb) and np.array_equal(va, vb), f"replay {k} differs"
    assert all(np.array_equal(p, q) for p, q in zip(a.contacts(), b.contacts()))


@pytest.mark.parametrize("name", ["tet_grid", "drop_on_box", "cloth_sphere", "sdf", "external", "springs"])
def test_graph_replays_equal_eager_steps(name):
    arr = configs.PARITY_SCENES[name]().arrays()
    a, b = pair(arr)
    n = len(arr["p_type"])
    check_steps(a, b, 8, ext=external_loads(n) if name == "external" else None)


def test_graph_recaptures_on_parameter_change_and_batches():
    arrs = [configs.PARITY_SCENES[n]().arrays() for n in ("drop_on_box", "cloth_sphere", "tet_grid")]
    a, b = pair(arrs)
    check_steps(a, b, 4, sub=20, tol=1e-5)
    check_steps(a, b, 4, sub=35, tol=1e-6)     # new parameters: a new graph
    b.load_state(a.save_state())                # checkpoint restore between replays
    check_steps(a, b, 3, sub=35, tol=1e-6)


def test_graph_with_precision_form():
    arr = configs.tet_grid(6, young_per_element=configs.log_uniform_young(6, seed=3)).arrays()
    a, b = pair(arr, mode=_lib.SOLVER_REFINED_BF16, inner_iterations=20, inner_tolerance=1e-3)
    check_steps(a, b, 5, sub=80, tol=1e-5)
