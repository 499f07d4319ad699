"""Fault statuses that replace the reference's asserts and silent divergence.

The reference asserts a positive Jacobi diagonal (sparse_solver.cpp:45-52) only in debug builds and has no check
for a diverging state. The device steps on as the reference's release build does, records the fault per instance
(b3g_step_stats::faults, sticky) and b3g_synchronize reports it once. The conditions are forced from the host -- a
NaN or infinite velocity, a negative Young's modulus -- so every kernel runs to its normal end.
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


def test_healthy_steps_raise_no_fault():
    b = Body(configs.tet_grid(4).arrays())
    for _ in range(3):
        b.Step(configs.DT, 1, 30, 0.0, 1e-5)
    b.synchronize()
    assert b.faults() == [0]


def test_non_finite_state_is_reported_once_and_named():
    arrs = [configs.tet_grid(3).arrays(), configs.tet_grid(4).arrays()]
    b = Body(arrs)
    b.Step(configs.DT, 1, 30, 0.0, 1e-5)
    b.synchronize()
    x, v = b.state()
    dyn = np.flatnonzero(np.concatenate([a["p_type"] for a in arrs]) == 2)
    v[dyn[-1], 0] = np.nan                     # a dynamic particle of instance 1
    b.set_state(v=v)
    b.Step(configs.DT, 1, 30, 0.0, 1e-5)
    with pytest.raises(_lib.B3gError, match=r"non-finite.*instance 1") as e:
        b.synchronize()
    assert e.value.status == _lib.B3G_NON_FINITE
    f = b.faults()
    assert f[0] == 0 and f[1] & _lib.FAULT_NON_FINITE
    b.synchronize()                             # reported once; the per-instance bits stay
    assert b.faults()[1] & _lib.FAULT_NON_FINITE


def test_non_positive_diagonal_is_reported():
    """A negative Young's modulus makes the element stiffness negative definite, so h^2 K outweighs the mass and
    the Jacobi diagonal of pA is negative (144 entries, as the oracle's pA shows): the reference would assert."""
    b = Body(configs.tet_grid(3, young=-1e7).arrays())
    b.Step(configs.DT, 1, 10, 0.0, 1e-5)
    with pytest.raises(_lib.B3gError, match="Jacobi diagonal") as e:
        b.synchronize()
    assert e.value.status == _lib.B3G_NON_POSITIVE_DIAGONAL
    assert b.faults()[0] & _lib.FAULT_DIAGONAL


def test_load_state_clears_faults():
    arr = configs.tet_grid(3).arrays()
    b = Body(arr)
    good = b.save_state()
    x, v = b.state()
    assert arr["p_type"][0] == 2      # a dynamic particle (SetVelocity ignores static ones, particle.h:225-230)
    v[0, 1] = np.inf
    b.set_state(v=v)
    b.Step(configs.DT, 1, 10, 0.0, 1e-5)
    with pytest.raises(_lib.B3gError):
        b.synchronize()
    b.load_state(good)
    assert b.faults() == [0]
    b.Step(configs.DT, 1, 10, 0.0, 1e-5)
    b.synchronize()
