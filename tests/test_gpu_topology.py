"""Topology edits (b3g_body_rebuild) and action forces (b3g_get_action_forces) against the reference.

The reference fixture tests/golden/tearing.npz was written by the compiled reference editing its live body
(b3Body::CreateParticle / CreateFixture / CreateForce / DestroyForce) between steps. The device rebuilds from the
edited description, taking over the previous body's state, and in the reference's summation order (sequential
reduction mode) continues bit for bit.
"""

import os

import numpy as np
import pytest

from bounce_softbody_amd import _lib, configs
from bounce_softbody_amd.body import Body
from bounce_softbody_amd.scene import edit_arrays, FORCE_SPRING, FORCE_STRETCH, FORCE_SHEAR, FORCE_MOUSE
import oracle
from tests.golden.make_golden import decode_edits, digest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "tearing.npz")


@pytest.fixture(scope="module", autouse=True)
def _device():
    if _lib.device_count() < 1:
        pytest.fail("no GPU visible to libb3g")


def test_tearing_matches_reference_golden():
    g = np.load(GOLDEN)
    arrays = configs.scene_tearing().arrays()
    b = Body(arrays)
    b.set_reduction(_lib.REDUCE_SEQUENTIAL)
    b.set_action_forces(True)
    pre, post = int(g["pre"]), int(g["post"])
    for k in range(pre):
        b.Step(configs.DT, 2, 20)
        assert np.array_equal(digest(*b.state()), g["digests"][k]), k
    assert np.array_equal(b.action_forces(FORCE_SPRING), g["action_at_edit"])
    x, v = b.state()
    edits = decode_edits(g["edits"])
    arr2, carry = edit_arrays(arrays, edits)
    b.rebuild(arr2, carry)
    for k in range(post):
        b.Step(configs.DT, 2, 20)
        assert np.array_equal(digest(*b.state()), g["digests"][pre + k]), k
    x, v = b.state()
    assert np.array_equal(x, g["x"]) and np.array_equal(v, g["v"])
    p, w, f = b.contacts()
    assert np.array_equal(p, g["contact_particle"]) and np.array_equal(w, g["contact_world"])
    assert np.array_equal(f, g["contact_force"])
    assert np.array_equal(b.action_forces(FORCE_SPRING), g["action_final"])


@pytest.mark.parametrize("name,ftype,per", [("springs", FORCE_SPRING, 2), ("cloth_shapes", FORCE_STRETCH, 3),
                                            ("cloth_shapes", FORCE_SHEAR, 3), ("cloth_sphere", FORCE_STRETCH, 3),
                                            ("mouse", FORCE_MOUSE, 4)])
def test_action_forces_equal_oracle(name, ftype, per):
    """m_f1.. summed over two Newton iterations per step equal the oracle's (pinned to the reference) bit for bit."""
    arr = configs.PARITY_SCENES[name]().arrays()
    b = Body(arr)
    b.set_reduction(_lib.REDUCE_SEQUENTIAL)
    b.set_action_forces(True)
    o = oracle.OracleBody(arr)
    for k in range(6):
        b.Step(configs.DT, 2, 20)
        o.step(configs.DT, 2, 20)
        assert np.array_equal(b.action_forces(ftype), o.action_forces(ftype, per)), k


def test_rebuild_carries_tets_and_contacts():
    """A rebuild with identity maps (no edit) changes nothing: the next steps equal those of an untouched body,
    tet warm-start quaternions and contacts included; without maps the contacts restart."""
    arr = configs.PARITY_SCENES["drop_on_box"]().arrays()
    a, b = Body(arr), Body(arr)
    for body in (a, b):
        for _ in range(10):
            body.Step(configs.DT, 1, 20, 0.0, 1e-5)
    assert len(a.contacts()[0]) > 0
    arr2, carry = edit_arrays(arr, [])
    b.rebuild(arr2, carry)
    assert np.array_equal(a.tet_quats(), b.tet_quats())
    assert all(np.array_equal(u, w) for u, w in zip(a.contacts(), b.contacts()))
    for _ in range(5):
        for body in (a, b):
            body.Step(configs.DT, 1, 20, 0.0, 1e-5)
        xa, va = a.state()
        xb, vb = b.state()
        assert np.array_equal(xa, xb) and np.array_equal(va, vb)
    c = Body(arr)
    for _ in range(3):
        c.Step(configs.DT, 1, 20, 0.0, 1e-5)
    c.rebuild(arr)
    assert len(c.contacts()[0]) == 0


def test_rebuild_validates_maps():
    arr = configs.PARITY_SCENES["drop_on_box"]().arrays()
    b = Body(arr)
    b.Step(configs.DT, 1, 20, 0.0, 1e-5)
    _, carry = edit_arrays(arr, [])
    x, v = b.state()
    carry["particle_prev"] = carry["particle_prev"] + 1000
    with pytest.raises(_lib.B3gError, match="carry map"):
        b.rebuild(arr, carry)
    x2, v2 = b.state()   # a rejected call changes nothing
    assert np.array_equal(x, x2) and np.array_equal(v, v2)
    b.Step(configs.DT, 1, 20, 0.0, 1e-5)


def test_batch_rebuild_with_identity_maps_changes_nothing():
    """b3g_body_rebuild on a batch: the maps index the instances' concatenated creation order. Identity maps over
    an uneven batch with contacts leave every later step equal to those of an untouched batch."""
    arrs = [configs.PARITY_SCENES[n]().arrays() for n in ("drop_on_box", "cloth_shapes", "springs", "tet_grid")]
    a, b = Body(arrs), Body(arrs)
    for body in (a, b):
        for _ in range(8):
            body.Step(configs.DT, 1, 20, 0.0, 1e-5)
    assert len(a.contacts()[0]) > 0
    total = lambda key, w=1: sum(len(x[key]) // w for x in arrs)
    carry = {"particle_prev": np.arange(total("p_type"), dtype=np.int64),
             "tet_prev": np.arange(total("tet_param", 3), dtype=np.int64),
             "sphere_prev": np.arange(total("s_particle"), dtype=np.int64),
             "world_prev": np.arange(total("w_type"), dtype=np.int64)}
    b.rebuild(arrs, carry)
    assert np.array_equal(a.tet_quats(), b.tet_quats())
    assert all(np.array_equal(u, w) for u, w in zip(a.contacts(), b.contacts()))
    for _ in range(5):
        for body in (a, b):
            body.Step(configs.DT, 1, 20, 0.0, 1e-5)
        xa, va = a.state()
        xb, vb = b.state()
        assert np.array_equal(xa, xb) and np.array_equal(va, vb)


def test_action_forces_cleared_by_zero_dt_step():
    """b3Body::Step clears the action forces before its dt > 0 test (body.cpp:563-574): after a step with dt = 0
    they read zero, as the oracle's do, and the next full step accumulates afresh."""
    arr = configs.PARITY_SCENES["springs"]().arrays()
    b = Body(arr)
    b.set_reduction(_lib.REDUCE_SEQUENTIAL)
    b.set_action_forces(True)
    o = oracle.OracleBody(arr)
    for dt in (configs.DT, configs.DT, 0.0, configs.DT):
        b.Step(dt, 2, 20)
        o.step(dt, 2, 20)
        got = b.action_forces(FORCE_SPRING)
        assert np.array_equal(got, o.action_forces(FORCE_SPRING, 2)), dt
        if dt == 0.0:
            assert not np.any(got)
        else:
            assert np.any(got)
