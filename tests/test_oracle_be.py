"""The oracle's b3SparseSolveBE (any force model, 3x3 diagonal blocks M and S, a sub-solution z) against the
reference's own b3SparseSolveBE, built from its sources (sparse_force_solver.cpp:104-232): bit for bit."""

import numpy as np
import pytest

import oracle
from be_models import problem

needs_ref = pytest.mark.skipif(not oracle.reference_available(), reason="reference build absent (oracle/_ref)")

CASES = {
    # name: (problem kwargs, maxIterations, tolerance, maxSubIterations, subTolerance)
    "plain": (dict(shape=(6, 5, 4), seed=1, general=False), 1, 1e-4, 50, 1e-5),
    "general": (dict(shape=(6, 5, 4), seed=2, general=True), 1, 1e-4, 50, 1e-5),
    "newton": (dict(shape=(5, 5, 5), seed=3, general=True), 4, 1e-6, 80, 1e-6),
    "no_damping": (dict(shape=(4, 6, 3), seed=4, general=True, damping=False), 2, 1e-5, 40, 1e-5),
    "capped": (dict(shape=(7, 6, 5), seed=5, general=True), 2, 1e-7, 7, 1e-9),
    "h_zero": (dict(shape=(4, 4, 4), seed=6, general=True, h=0.0), 2, 1e-4, 30, 1e-5),
    "hh_underflow": (dict(shape=(4, 4, 4), seed=7, general=True, h=1e-23), 1, 1e-4, 30, 1e-5),
    "no_newton": (dict(shape=(3, 3, 3), seed=8, general=True), 0, 1e-4, 30, 1e-5),
    # springs add Jacobian blocks only while stretched: the pattern changes between Newton iterations
    "changing_pattern": (dict(shape=(6, 6, 4), seed=9, general=True, tension_only=True), 5, 1e-7, 60, 1e-6),
}


def run(which, p, it, tol, sub, subtol, minmax=(0xFFFFFFFF, 0), wide=False):
    return oracle.solve_be(which, p["model"], p["h"], p["x0"], p["v0"], p["fe"], p["M"], p["y"], p["S"], p["z"],
                           it, tol, sub, subtol, minmax, wide=wide)


@needs_ref
@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_solve_be_equals_reference(name):
    kw, it, tol, sub, subtol = CASES[name]
    ro = run("oracle", problem(**kw), it, tol, sub, subtol)
    rr = run("ref", problem(**kw), it, tol, sub, subtol)
    assert np.array_equal(ro[0], rr[0]) and np.array_equal(ro[1], rr[1])
    assert ro[2:] == rr[2:]
    if it:
        assert np.isfinite(ro[0]).all() and np.abs(ro[1] - problem(**kw)["v0"]).max() > 0


@needs_ref
def test_running_min_max_sub_iterations_fold_into_the_callers_values():
    """The output's minSubIterations / maxSubIterations keep the values the caller put there
    (sparse_force_solver.cpp:186-188)."""
    kw, it, tol, sub, subtol = CASES["newton"]
    for mm in ((0xFFFFFFFF, 0), (3, 5), (0, 1000)):
        ro = run("oracle", problem(**kw), it, tol, sub, subtol, mm)
        rr = run("ref", problem(**kw), it, tol, sub, subtol, mm)
        assert ro[4] == rr[4]
        assert ro[4][0] <= mm[0] and ro[4][1] >= mm[1]


def test_free_particle_closed_form():
    """No force Jacobians, S = I, M = m I, no forces but gravity: one Newton iteration solves the linear system
    M dv = h fe exactly enough that v = v0 + h g and x = x0 + h v + y (sparse_force_solver.cpp:60-90)."""
    n = 5
    rng = np.random.default_rng(0)
    m = rng.uniform(0.5, 2.0, n).astype(np.float32)
    g = np.array([0.0, -9.8, 0.0], np.float32)
    h = np.float32(0.01)
    eye = np.eye(3, dtype=np.float32).reshape(9)
    M = (m[:, None] * eye).astype(np.float32)
    S = np.tile(eye, (n, 1))
    x0 = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    v0 = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    fe = (m[:, None] * g).astype(np.float32)
    y = np.zeros((n, 3), np.float32)

    def model(x, v):
        return np.zeros((len(x), 3), np.float32), []

    x, v, it, err, mm = oracle.solve_be("oracle", model, h, x0, v0, fe, M, y, S, None, 1, 1e-4, 10, 1e-6)
    np.testing.assert_allclose(v, v0 + h * g, rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(x, x0 + h * v, rtol=1e-6, atol=1e-6)
    assert it == 1   # |dv|^2 > tol^2: the loop ends at maxIterations, not on convergence
