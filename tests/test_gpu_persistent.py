"""The persistent PCG (schedule persist_tiles / persist_groups): the instances of at most persist_tiles SpMV tiles
run all their PCG iterations in one launch, one workgroup per instance solve, beside the launch sequence of the
other instances. Its tiles, row blocks and FP64 partial sums are the launch sequence's (sparse_solver.cpp:55-103
per row and block), so every step is bit-identical to the default schedule's: states, statistics, contacts."""

import numpy as np
import pytest

from bounce_softbody_amd import _lib, configs
from bounce_softbody_amd.body import Body
from tests.golden.make_golden import external_loads

pytestmark = pytest.mark.gpu

SCENES = ["tet_grid", "node_types", "drop_on_box", "cloth_sphere", "cloth_shapes", "springs", "sdf", "external",
          "mesh_terrain"]


@pytest.fixture(scope="module", autouse=True)
def _device():
    if _lib.device_count() < 1:
        pytest.fail("no GPU visible to libb3g")


def uneven_batch():
    arrs = [configs.PARITY_SCENES[n]().arrays() for n in SCENES]
    return arrs + [configs.tet_grid(k).arrays() for k in (2, 7, 4, 12)]


def same(a, b):
    xa, va = a.state()
    xb, vb = b.state()
    sa, sb = a.instance_stats(), b.instance_stats()
    return (np.array_equal(xa, xb) and np.array_equal(va, vb)
            and all(np.array_equal(p[0], q[0]) and p[1] == q[1] for p, q in zip(sa, sb))
            and all(np.array_equal(p, q) for p, q in zip(a.contacts(), b.contacts())))


def step_both(ref, b, steps, arrs, it=2, sub=40, stol=1e-5):
    n = sum(len(a["p_type"]) for a in arrs)
    for k in range(steps):
        if k % 4 == 1:
            f, t = external_loads(n)
            ref.apply_external(f, t)
            b.apply_external(f, t)
        ref.Step(configs.DT, it, sub, 0.0, stol)
        b.Step(configs.DT, it, sub, 0.0, stol)
        assert same(ref, b), f"step {k}"


@pytest.mark.parametrize("persist", [(1 << 20, 0), (6, 0), (1 << 20, 3), (2, 1)])
@pytest.mark.parametrize("mode", [_lib.REDUCE_FP64, _lib.REDUCE_FP64_SEPARATE])
def test_persistent_instances_equal_the_launch_sequence(persist, mode):
    print(f"persistent {persist} mode {mode}", flush=True)
    arrs = uneven_batch()
    ref, b = Body(arrs), Body(arrs)
    for x in (ref, b):
        x.set_reduction(mode)
    b.set_schedule(*_lib.DEFAULT_SCHEDULE[:5], *persist)
    assert b.schedule()[5:] == persist
    step_both(ref, b, 8, arrs)


def test_persistent_with_captured_steps_and_without_early_exit():
    """Captured steps record the fork to the second stream and the join; replays equal eager launches."""
    arrs = uneven_batch()
    ref, b = Body(arrs), Body(arrs)
    b.set_schedule(*_lib.DEFAULT_SCHEDULE[:5], 1 << 20, 4)
    b.set_graph(True)
    step_both(ref, b, 6, arrs)
    b.set_graph(False)
    b.set_early_exit(False)
    step_both(ref, b, 3, arrs)


def test_persistent_single_body_and_sequential_mode():
    """A single body can run persistent; the sequential reduction mode keeps the launches (bit-exact either way)."""
    arrs = configs.tet_grid(6).arrays()
    ref, b = Body(arrs), Body(arrs)
    b.set_schedule(*_lib.DEFAULT_SCHEDULE[:5], 1 << 20, 0)
    step_both(ref, b, 4, [arrs], it=1, sub=50)
    for x in (ref, b):
        x.set_reduction(_lib.REDUCE_SEQUENTIAL)
    step_both(ref, b, 3, [arrs], it=1, sub=50)


def test_persist_tiles_limit_is_checked():
    b = Body(configs.tet_grid(2).arrays())
    with pytest.raises(_lib.B3gError) as e:
        b.set_schedule(*_lib.DEFAULT_SCHEDULE[:5], (1 << 20) + 1, 0)
    assert e.value.status == _lib.B3G_INVALID_ARGUMENT


def test_box_batch_persistent_equals_launches():
    """Bodies of the many-body batch (3..24 cells per axis), captured steps: every instance persistent, and the
    instances of at most 24 or 16 tiles persistent beside the launches of the others (some on fewer workgroups than
    instances), bit-identical to the default launches. The mixed runs catch partial sums shared between the
    persistent instances and the launch sequence, which run out of step."""
    arrs = configs.batch_boxes(512)
    for persist in ((1 << 20, 0), (24, 0), (16, 128)):
        print(f"box batch persistent {persist}", flush=True)
        ref, b = Body(arrs), Body(arrs)
        for x in (ref, b):
            x.set_graph(True)
        b.set_schedule(*_lib.DEFAULT_SCHEDULE[:5], *persist)
        step_both(ref, b, 5, arrs, it=1, sub=50)
