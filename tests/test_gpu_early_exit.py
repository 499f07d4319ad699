"""Early exit of the PCG (b3g_set_early_exit): the reference leaves its PCG loop once the residual test passes
(sparse_solver.cpp:63-73). The device enqueues the iterations in chunks, and once no instance's last iteration
ran it stops enqueuing the rest, which would all have returned at once. So every state, statistic and contact
equals the fixed launch sequence's bit for bit: single bodies, uneven batches, contact scenes, several Newton
iterations, the refined solver and captured steps.
"""

import numpy as np
import pytest

from bounce_softbody_amd import _lib, configs
from bounce_softbody_amd.body import Body

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    if _lib.device_count() < 1:
        pytest.fail("no GPU visible to libb3g")


def same(a, b):
    xa, va = a.state()
    xb, vb = b.state()
    return (np.array_equal(xa, xb) and np.array_equal(va, vb) and
            all(np.array_equal(sa[0], sb[0]) and sa[1] == sb[1] for sa, sb in zip(a.instance_stats(), b.instance_stats()))
            and all(np.array_equal(p, q) for p, q in zip(a.contacts(), b.contacts())))


def run_pair(arrs, steps, schedule, setup=None):
    on, off = Body(arrs), Body(arrs)
    off.set_early_exit(False)
    for b in (on, off):
        if setup:
            setup(b)
    for k in range(steps):
        it, sub, stol = schedule(k)
        on.Step(configs.DT, it, sub, 0.0, stol)
        off.Step(configs.DT, it, sub, 0.0, stol)
        assert same(on, off), f"step {k}"
    return on, off


@pytest.mark.parametrize("name", ["tet_grid", "drop_on_box", "cloth_sphere", "springs", "sdf", "mesh_terrain"])
def test_early_exit_equals_fixed_sequence(name):
    arr = configs.PARITY_SCENES[name]().arrays()
    # converged solves (tolerance 1e-5) and capped ones (tolerance 1e-9), one and two Newton iterations
    run_pair(arr, 12, lambda k: (1 + k % 2, 40, 1e-5 if k % 3 else 1e-9))


def test_early_exit_uneven_batch():
    arrs = [configs.PARITY_SCENES[n]().arrays() for n in ("tet_grid", "drop_on_box", "cloth_shapes", "springs")]
    arrs += [configs.tet_grid(k).arrays() for k in (2, 6, 3)]
    run_pair(arrs, 10, lambda k: (1, 50, 1e-5))


def test_early_exit_refined_solver():
    arr = configs.tet_grid(8, young_per_element=configs.log_uniform_young(8, seed=7)).arrays()
    run_pair(arr, 6, lambda k: (1, 200, 1e-6), setup=lambda b: b.set_solver(_lib.SOLVER_REFINED_FP32, 30, 1e-3))


def test_early_exit_with_sequential_reduction():
    arr = configs.PARITY_SCENES["cloth_shapes"]().arrays()
    run_pair(arr, 8, lambda k: (2, 20, 1e-5), setup=lambda b: b.set_reduction(_lib.REDUCE_SEQUENTIAL))


def test_captured_steps_unaffected():
    """Captured steps record the fixed sequence; with early exit on for the eager body, replays still equal it."""
    arr = configs.PARITY_SCENES["drop_on_box"]().arrays()
    eager, captured = Body(arr), Body(arr)
    captured.set_graph(True)
    for k in range(8):
        eager.Step(configs.DT, 1, 30, 0.0, 1e-5)
        captured.Step(configs.DT, 1, 30, 0.0, 1e-5)
        assert same(eager, captured), k


def test_early_exit_full_size_tet():
    """The 64^3 tet grid of the benchmark: 4 steps with and without early exit, bit for bit."""
    run_pair(configs.tet_grid(64).arrays(), 4, lambda k: (1, 50, 1e-5))
