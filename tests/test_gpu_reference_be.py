"""The operator seam as an integration would use it: the reference's own library, built from its sources with
sparse/sparse_force_solver.cpp replaced by tests/native/b3g_sparse_solve_be.cpp (oracle/Makefile, target
ref_be). Its b3Body::Step, force elements and contacts run unchanged on the host; every b3SparseSolveBE call
(force_solver.cpp:165) goes to b3g_solve_be on the GPU. In sequential reduction mode the goldens the
unmodified reference wrote are reproduced bit for bit."""

import os

import numpy as np
import pytest

import oracle
from bounce_softbody_amd import _lib
from tests.golden.make_golden import digest
from tests.test_oracle_golden import FIXTURES, GOLDEN, replay, scene_arrays

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(oracle.REF_BE_SO), reason="oracle/_ref/libb3ref_be.so not built")]


class RefBodyBE(oracle.RefBody):
    which = "ref_be"


def set_mode(mode):
    oracle._lib("ref_be").refh_be_set_reduction(mode)


def last_status():
    return oracle._lib("ref_be").refh_be_last_status()


@pytest.mark.parametrize("name", FIXTURES)
def test_reference_step_with_the_device_solve_reproduces_the_golden(name):
    print(f"ref_be sequential {name}", flush=True)
    set_mode(_lib.REDUCE_SEQUENTIAL)
    g = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    arr = scene_arrays(name)
    body = RefBodyBE(arr)
    n = len(arr["p_type"])
    for k in replay(body, g, n):
        assert last_status() == _lib.B3G_OK
        assert np.array_equal(digest(*body.state()), g["digests"][k]), f"{name}: state differs after step {k}"
    x, v = body.state()
    assert np.array_equal(x, g["x"]) and np.array_equal(v, g["v"])
    cp, cw, cf = body.contacts()
    assert np.array_equal(cp, g["contact_particle"]) and np.array_equal(cw, g["contact_world"])
    assert np.array_equal(cf, g["contact_normal_force"])
    assert np.array_equal(body.tet_quats(), g["quats"])
    assert np.array_equal(body.stats(), g["stats"])


@pytest.mark.parametrize("name", ["tet_grid", "cloth_sphere", "springs"])
def test_reference_step_with_the_fp64_device_solve_stays_within_1e4(name):
    """Default reduction (FP64 sums): each of the first ten steps of the golden schedule, started from the
    unmodified reference's state, lands within 1e-4 relative max-norm of the reference's step."""
    print(f"ref_be fp64 {name}", flush=True)
    set_mode(_lib.REDUCE_FP64)
    try:
        g = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
        arr = scene_arrays(name)
        dev, ref = RefBodyBE(arr), oracle.RefBody(arr)
        n = len(arr["p_type"])
        for k, _ in zip(replay(dev, g, n), replay(ref, g, n)):
            (dx, dv), (rx, rv) = dev.state(), ref.state()
            ex = float(np.abs(dx - rx).max() / max(np.abs(rx).max(), 1e-30))
            ev = float(np.abs(dv - rv).max() / max(np.abs(rv).max(), 1e-30))
            assert ex <= 1e-4 and ev <= 1e-4, (name, k, ex, ev)
            if k == 9:
                break
            dev.set_state(rx, rv)   # the next step starts from the reference's state
    finally:
        set_mode(_lib.REDUCE_SEQUENTIAL)


def test_reference_step_with_the_device_solve_at_full_size():
    """The 64^3 tet grid (274,625 particles, 1.31 M tetrahedra): one step of the reference's own b3Body::Step with
    b3SparseSolveBE on the GPU, PCG <= 50 at 1e-5, against the unmodified reference from the same set-up --
    bit for bit in sequential mode, within 1e-4 in FP64 mode."""
    import time
    from bounce_softbody_amd import configs
    arr = configs.tet_grid(64).arrays()
    t0 = time.perf_counter()
    ref = oracle.RefBody(arr, fast=True)
    ref.step(configs.DT, 1, 50, 0.0, 1e-5)
    t_ref = time.perf_counter() - t0
    rx, rv = ref.state()
    for mode in (_lib.REDUCE_SEQUENTIAL, _lib.REDUCE_FP64):
        set_mode(mode)
        try:
            t0 = time.perf_counter()
            dev = RefBodyBE(arr, fast=True)
            dev.step(configs.DT, 1, 50, 0.0, 1e-5)
            t_dev = time.perf_counter() - t0
            assert last_status() == _lib.B3G_OK
            dx, dv = dev.state()
            print(f"64^3 mode {mode}: reference set-up + step {t_ref:.1f} s, with the device solve {t_dev:.1f} s",
                  flush=True)
            if mode == _lib.REDUCE_SEQUENTIAL:
                assert np.array_equal(dx, rx) and np.array_equal(dv, rv)
                assert np.array_equal(dev.stats(), ref.stats())
            else:
                assert np.abs(dx - rx).max() <= 1e-4 * np.abs(rx).max()
                assert np.abs(dv - rv).max() <= 1e-4 * np.abs(rv).max()
            del dev
        finally:
            set_mode(_lib.REDUCE_SEQUENTIAL)
