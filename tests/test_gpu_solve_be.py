"""b3g_solve_be -- b3SparseSolveBE behind a host force model (sparse_force_solver.cpp:104-232) -- on the GPU,
through the C ABI, against the oracle (itself pinned bit for bit to the reference's b3SparseSolveBE by
tests/test_oracle_be.py)."""

import numpy as np
import pytest

import oracle
from bounce_softbody_amd import _lib
from bounce_softbody_amd.sparse import BESolver, SolveBEInput, SolveBEOutput
from be_models import problem
from test_oracle_be import CASES

pytestmark = pytest.mark.gpu


def device_model(model):
    """The (f, ops) model as a b3SparseForceModel callback."""
    def apply(data):
        f, ops = model(data.x, data.v)
        data.f[:] = f
        for target, sign, rows, cols, blocks in ops:
            m = data.dfdx if target == "x" else data.dfdv
            (m.add if sign > 0 else m.sub)(rows, cols, blocks)
    return apply


def gpu(solver, p, it, tol, sub, subtol, minmax=(0xFFFFFFFF, 0)):
    inp = SolveBEInput(h=float(p["h"]), inv_h=0.0 if p["h"] == 0 else float(np.float32(1) / p["h"]),
                       forceModel=device_model(p["model"]), x0=p["x0"], v0=p["v0"], fe=p["fe"], M=p["M"], y=p["y"],
                       S=p["S"], z=p["z"], maxIterations=it, tolerance=tol, maxSubIterations=sub, subTolerance=subtol)
    out = SolveBEOutput(minSubIterations=minmax[0], maxSubIterations=minmax[1])
    solver.solve(out, inp)
    return out.x, out.v, out.iterations, float(np.float32(out.error)), (out.minSubIterations, out.maxSubIterations)


def cpu(p, it, tol, sub, subtol, minmax=(0xFFFFFFFF, 0), wide=False):
    return oracle.solve_be("oracle", p["model"], p["h"], p["x0"], p["v0"], p["fe"], p["M"], p["y"], p["S"], p["z"],
                           it, tol, sub, subtol, minmax, wide=wide)


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


@pytest.mark.parametrize("name", sorted(CASES))
def test_sequential_mode_equals_the_oracle_bit_for_bit(name):
    kw, it, tol, sub, subtol = CASES[name]
    print(f"solve_be sequential {name}", flush=True)
    s = BESolver(reduction=_lib.REDUCE_SEQUENTIAL)
    g = gpu(s, problem(**kw), it, tol, sub, subtol)
    o = cpu(problem(**kw), it, tol, sub, subtol)
    assert np.array_equal(g[0], o[0]) and np.array_equal(g[1], o[1]), (rel(g[0], o[0]), rel(g[1], o[1]))
    assert g[2:] == o[2:]


@pytest.mark.parametrize("name", sorted(CASES))
def test_fp64_mode_against_the_widened_and_the_reference_accumulation(name):
    """Default reduction: the PCG dot products and the Newton error summed in FP64, as b3g_step sums them.
    Against the oracle with its accumulators widened the same way: within 1e-6; against the reference's FP32
    accumulation: within 1e-4 relative max-norm, Newton counts equal, PCG counts within one."""
    kw, it, tol, sub, subtol = CASES[name]
    print(f"solve_be fp64 {name}", flush=True)
    s = BESolver()
    g = gpu(s, problem(**kw), it, tol, sub, subtol)
    w = cpu(problem(**kw), it, tol, sub, subtol, wide=True)
    o = cpu(problem(**kw), it, tol, sub, subtol)
    assert rel(g[0], w[0]) <= 1e-6 and rel(g[1], w[1]) <= 1e-6
    assert rel(g[0], o[0]) <= 1e-4 and rel(g[1], o[1]) <= 1e-4
    assert g[2] == o[2]
    assert abs(g[4][0] - o[4][0]) <= 1 and abs(g[4][1] - o[4][1]) <= 1


def test_one_solver_across_patterns_sizes_and_running_counts():
    """The context keeps its buffers and symbolic pattern between calls and rebuilds them when the force
    model's pattern or the size changes; the caller's running min / max PCG counts are folded into."""
    s = BESolver(reduction=_lib.REDUCE_SEQUENTIAL)
    seq = [("general", (0xFFFFFFFF, 0)), ("newton", (3, 5)), ("general", (0, 1000)), ("h_zero", (0xFFFFFFFF, 0)),
           ("plain", (7, 7)), ("plain", (0xFFFFFFFF, 0))]
    for name, mm in seq:
        kw, it, tol, sub, subtol = CASES[name]
        print(f"solve_be reuse {name} {mm}", flush=True)
        g = gpu(s, problem(**kw), it, tol, sub, subtol, mm)
        o = cpu(problem(**kw), it, tol, sub, subtol, mm)
        assert np.array_equal(g[0], o[0]) and np.array_equal(g[1], o[1]) and g[2:] == o[2:], name


def test_force_model_is_called_once_per_newton_iteration():
    kw, it, tol, sub, subtol = CASES["newton"]
    p = problem(**kw)
    g = gpu(BESolver(reduction=_lib.REDUCE_SEQUENTIAL), p, it, tol, sub, subtol)
    calls = p["model"].calls
    q = problem(**kw)
    o = cpu(q, it, tol, sub, subtol)
    assert calls == q["model"].calls == min(it, g[2] + 1)
    assert g[2] == o[2]


def test_force_model_errors_end_the_solve():
    kw, it, tol, sub, subtol = CASES["general"]
    p = problem(**kw)

    class Boom(RuntimeError):
        pass

    def bad(data):
        raise Boom("model failed")

    inp = SolveBEInput(h=float(p["h"]), inv_h=60.0, forceModel=bad, x0=p["x0"], v0=p["v0"], fe=p["fe"], M=p["M"],
                       y=p["y"], S=p["S"], z=p["z"], maxIterations=1, maxSubIterations=20)
    with pytest.raises(Boom):
        BESolver().solve(SolveBEOutput(), inp)

    def out_of_range(data):
        data.dfdx.add([len(data.x)], [0], np.zeros(9, np.float32))

    inp.forceModel = out_of_range
    with pytest.raises(_lib.B3gError, match="outside"):
        BESolver().solve(SolveBEOutput(), inp)


def test_non_positive_diagonal_is_reported():
    """A negative mass block makes pA's Jacobi diagonal negative (the reference asserts, sparse_solver.cpp:45-52,
    and its release build goes on): the solve completes as the release build would and reports the fault."""
    kw, it, tol, sub, subtol = CASES["plain"]
    p = problem(**kw)
    p["M"] = -p["M"]
    with pytest.raises(_lib.B3gError) as e:
        gpu(BESolver(), p, 1, tol, 5, subtol)
    assert e.value.status == _lib.B3G_NON_POSITIVE_DIAGONAL


def test_large_network_fp64_mode():
    """A 40^3-particle spring network (64,000 rows, ~1.6 M blocks of pA), two Newton iterations."""
    kw = dict(shape=(40, 40, 40), seed=11, general=True)
    print("solve_be 40^3", flush=True)
    g = gpu(BESolver(), problem(**kw), 2, 1e-6, 50, 1e-5)
    w = cpu(problem(**kw), 2, 1e-6, 50, 1e-5, wide=True)
    assert rel(g[0], w[0]) <= 1e-6 and rel(g[1], w[1]) <= 1e-6
    assert g[2] == w[2] and abs(g[4][1] - w[4][1]) <= 1


class ListMatrix:
    """The reference's b3SparseMat33 list semantics in Python (a row's head is the end of its list):
    operator()(i, j) searches from the head and inserts a zero block there when absent."""

    def __init__(self, n):
        self.rows = [[] for _ in range(n)]

    def apply(self, sign, rows, cols, blocks):
        for i, j, K in zip(rows, cols, np.asarray(blocks, np.float32).reshape(-1, 9)):
            r = self.rows[int(i)]
            for e in reversed(r):
                if e[0] == j:
                    e[1] = (e[1] + K if sign > 0 else e[1] - K).astype(np.float32)
                    break
            else:
                z = np.zeros(9, np.float32)
                r.append([int(j), (z + K if sign > 0 else z - K).astype(np.float32)])


def test_rows_handed_over_with_set_row_equal_incremental_adds():
    """An integration that lets the reference's own b3SparseMat33 collect the force model's blocks hands each
    row over head to tail with b3g_matrix_set_row: the solve equals the one whose model adds block by block."""
    kw, it, tol, sub, subtol = CASES["newton"]
    p = problem(**kw)
    model = p["model"]

    def by_rows(data):
        f, ops = model(data.x, data.v)
        data.f[:] = f
        mats = {"x": ListMatrix(len(f)), "v": ListMatrix(len(f))}
        for target, sign, rows, cols, blocks in ops:
            mats[target].apply(sign, rows, cols, blocks)
        for target, dst in (("x", data.dfdx), ("v", data.dfdv)):
            for i, r in enumerate(mats[target].rows):
                if r:
                    head_to_tail = r[::-1]
                    dst.set_row(i, [e[0] for e in head_to_tail], np.stack([e[1] for e in head_to_tail]))

    s = BESolver(reduction=_lib.REDUCE_SEQUENTIAL)
    inp = SolveBEInput(h=float(p["h"]), inv_h=float(np.float32(1) / p["h"]), forceModel=by_rows, x0=p["x0"],
                       v0=p["v0"], fe=p["fe"], M=p["M"], y=p["y"], S=p["S"], z=p["z"], maxIterations=it,
                       tolerance=tol, maxSubIterations=sub, subTolerance=subtol)
    out = s.solve(SolveBEOutput(), inp)
    o = cpu(problem(**kw), it, tol, sub, subtol)
    assert np.array_equal(out.x, o[0]) and np.array_equal(out.v, o[1]) and out.iterations == o[2]
