"""The PCG's launch schedule (b3g_get_schedule / b3g_set_schedule) and its autotuner (b3g_autotune).

A schedule groups rows into SpMV tiles and vector row blocks and picks where the FP64 dot-product partials are
summed; every row and block is computed as the reference computes it (sparse_solver.cpp:55-100), so steps under any
schedule agree with the default's to FP32 rounding of the FP64 sums. Changing the schedule keeps the assembled
system and the step statistics, and the autotuner's choice goes through an on-disk cache keyed by the system's shape.
"""

import numpy as np
import pytest

from bounce_softbody_amd import _lib, configs
from bounce_softbody_amd.body import Body

pytestmark = pytest.mark.gpu

SCHEDULES = [(1024, 256, 1024, 0, 0), (512, 128, 256, 1, 0), (32, 16, 64, 0, 256), (256, 64, 512, 1, 0),
             (1024, 256, 64, -1, 300)]


@pytest.fixture(scope="module", autouse=True)
def _device():
    if _lib.device_count() < 1:
        pytest.fail("no GPU visible to libb3g")


def scene():
    arrs = [configs.PARITY_SCENES[n]().arrays() for n in ("tet_grid", "drop_on_box", "cloth_shapes")]
    return arrs + [configs.tet_grid(k).arrays() for k in (2, 9, 5)]


def close(a, b, tol=1e-5):
    xa, va = a.state()
    xb, vb = b.state()
    return np.allclose(xa, xb, rtol=0, atol=tol) and np.allclose(va, vb, rtol=0, atol=60 * tol)


def test_default_schedule_and_validation():
    b = Body(configs.tet_grid(3).arrays())
    assert b.schedule() == _lib.DEFAULT_SCHEDULE
    for bad in [(16, 256, 1024, -1), (1025, 256, 1024, -1), (1024, 8, 1024, -1), (1024, 257, 1024, -1),
                (1024, 256, 32, -1), (1024, 256, 2048, -1), (1024, 256, 1024, 2), (1024, 256, 1024, -1, 100)]:
        with pytest.raises(_lib.B3gError) as e:
            b.set_schedule(*bad)
        assert e.value.status == _lib.B3G_INVALID_ARGUMENT
    assert b.schedule() == _lib.DEFAULT_SCHEDULE
    b.set_schedule(512, 128, 256, 1, 4096)
    assert b.schedule() == (512, 128, 256, 1, 4096, 0, 0)


@pytest.mark.parametrize("sched", SCHEDULES)
def test_steps_under_a_schedule_match_the_default(sched):
    arrs = scene()
    ref, b = Body(arrs), Body(arrs)
    b.set_schedule(*sched)
    for k in range(6):
        ref.Step(configs.DT, 1 + k % 2, 40, 0.0, 1e-5)
        b.Step(configs.DT, 1 + k % 2, 40, 0.0, 1e-5)
        assert close(ref, b), f"{sched}: step {k}"


@pytest.mark.parametrize("fused", [0, 1])
def test_ranges_per_workgroup_are_bit_identical(fused):
    """max_groups below the item count gives each PCG workgroup a range of tiles, skipping converged instances 64
    at a time: the same tiles and partials as one workgroup per tile, so the same bits."""
    # more than 512 tiles of 16 rows and row blocks of 64 rows, so both kernel kinds take ranges
    arrs = [configs.tet_grid(k).arrays() for k in (10, 3, 32, 6, 10, 4, 8, 7)]
    arrs += [configs.PARITY_SCENES[n]().arrays() for n in ("drop_on_box", "cloth_shapes")]
    assert sum(len(a["p_type"]) for a in arrs) // 64 >= 512
    one, ranged = Body(arrs), Body(arrs)
    one.set_schedule(1024, 16, 64, fused, 0)
    ranged.set_schedule(1024, 16, 64, fused, 256)
    for k in range(6):
        it, stol = (1 + k % 2, 1e-5 if k % 3 else 1e-9)
        one.Step(configs.DT, it, 40, 0.0, stol)
        ranged.Step(configs.DT, it, 40, 0.0, stol)
        xa, va = one.state()
        xb, vb = ranged.state()
        assert np.array_equal(xa, xb) and np.array_equal(va, vb), k
        assert all(np.array_equal(p[0], q[0]) for p, q in zip(one.instance_stats(), ranged.instance_stats())), k


def test_set_schedule_keeps_the_system_and_stats():
    arrs = scene()
    b = Body(arrs)
    b.Step(configs.DT, 1, 30, 0.0, 1e-5)
    b.synchronize()
    before = b.system()
    stats = b.instance_stats()
    b.set_schedule(32, 16, 64, 1)
    after = b.system()
    assert all(np.array_equal(p, q) for p, q in zip(before, after))
    assert all(np.array_equal(s[0], t[0]) and s[1] == t[1] for s, t in zip(stats, b.instance_stats()))


def test_schedule_with_captured_steps_and_sequential_reduction():
    arr = configs.PARITY_SCENES["drop_on_box"]().arrays()
    eager, captured = Body(arr), Body(arr)
    captured.set_graph(True)
    for b in (eager, captured):
        b.set_schedule(256, 64, 128, 0)
    for k in range(6):
        eager.Step(configs.DT, 1, 30, 0.0, 1e-5)
        captured.Step(configs.DT, 1, 30, 0.0, 1e-5)
        xe, ve = eager.state()
        xc, vc = captured.state()
        assert np.array_equal(xe, xc) and np.array_equal(ve, vc), k
    # the sequential reduction sums in row order whatever the tiles: bit for bit with the default schedule
    ref, b = Body(arr), Body(arr)
    for x in (ref, b):
        x.set_reduction(_lib.REDUCE_SEQUENTIAL)
    b.set_schedule(64, 32, 64, 1)
    for k in range(6):
        ref.Step(configs.DT, 1, 30, 0.0, 1e-5)
        b.Step(configs.DT, 1, 30, 0.0, 1e-5)
        xr, vr = ref.state()
        xb, vb = b.state()
        assert np.array_equal(xr, xb) and np.array_equal(vr, vb), k


def test_autotune_keeps_state_and_caches_its_choice(tmp_path):
    cache = str(tmp_path / "schedules.txt")
    arrs = scene()
    ref, b = Body(arrs), Body(arrs)
    for x in (ref, b):
        x.Step(configs.DT, 1, 40, 0.0, 1e-5)
    b.synchronize()
    before, x0, v0, stats, contacts = b.system(), *b.state(), b.instance_stats(), b.contacts()
    chosen, tuned_ms, default_ms = b.autotune(cache)   # times whole steps from this state, then restores it
    assert 0 < tuned_ms <= default_ms
    assert b.schedule() == chosen
    tb, tr, rb, fu, mg, pt, pg = chosen
    assert 32 <= tb <= 1024 and 16 <= tr <= 256 and 64 <= rb <= 1024 and fu in (-1, 0, 1) and (mg == 0 or mg >= 256)
    assert pt <= 1 << 20
    assert all(np.array_equal(p, q) for p, q in zip(before, b.system()))
    x1, v1 = b.state()
    assert np.array_equal(x0, x1) and np.array_equal(v0, v1)
    assert all(np.array_equal(s[0], t[0]) and s[1] == t[1] for s, t in zip(stats, b.instance_stats()))
    assert all(np.array_equal(p, q) for p, q in zip(contacts, b.contacts()))
    for k in range(4):
        ref.Step(configs.DT, 1, 40, 0.0, 1e-5)
        b.Step(configs.DT, 1, 40, 0.0, 1e-5)
        assert close(ref, b), f"step {k}"
    lines = open(cache).read().splitlines()
    assert len(lines) == 1
    # a body of the same shape hits the cache: same schedule, the recorded times, no search
    c = Body(arrs)
    hit, tm, dm = c.autotune(cache)
    assert hit == chosen and abs(tm - tuned_ms) < 1e-5 and abs(dm - default_ms) < 1e-5
    assert len(open(cache).read().splitlines()) == 1
    # another shape misses and appends its own line
    d = Body(configs.tet_grid(6).arrays())
    d.autotune(cache)
    assert len(open(cache).read().splitlines()) == 2


def test_autotune_captured_body_continues_like_an_untuned_one():
    arr = configs.PARITY_SCENES["cloth_sphere"]().arrays()
    ref, b = Body(arr), Body(arr)
    for x in (ref, b):
        x.set_graph(True)
        x.Step(configs.DT, 1, 30, 0.0, 1e-5)
    b.autotune()
    for k in range(4):
        ref.Step(configs.DT, 1, 30, 0.0, 1e-5)
        b.Step(configs.DT, 1, 30, 0.0, 1e-5)
        assert close(ref, b), f"step {k}"


def test_autotune_before_any_step():
    """A body that never stepped tunes on its zero system; its first steps are unaffected."""
    arr = configs.tet_grid(10).arrays()
    ref, b = Body(arr), Body(arr)
    b.autotune()
    for k in range(3):
        ref.Step(configs.DT, 1, 40, 0.0, 1e-5)
        b.Step(configs.DT, 1, 40, 0.0, 1e-5)
        assert close(ref, b), f"step {k}"
