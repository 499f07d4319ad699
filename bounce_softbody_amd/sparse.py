"""Host-side mirror of the reference's operator API for the implicit solve.

``SparseSolveBE(output, input)`` stands for ``b3SparseSolveBE``
(sparse/sparse_force_solver.h:97, sparse_force_solver.cpp:104-232), and a
:class:`SparseForceModel` for its plugin ``b3SparseForceModel``
(sparse_force_solver.h:44-49). The force model runs here, on the host, once
per Newton iteration; the linearisation, the pre-filtered PCG and the Newton
update run on the GPU through ``b3g_solve_be`` (include/b3g.h).
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import B3gError, _ptr, check, f32p, u32p

FLT_EPSILON = float(np.finfo(np.float32).eps)


class SparseMat33:
    """The ``b3SparseMat33`` a force model fills (a ``b3g_matrix``): ``add`` is ``A(i, j) += K`` and ``sub``
    is ``A(i, j) -= K`` for each block in order (sparse_mat33.h:268-288), blocks column-major (k, 3, 3) or
    (k, 9)."""

    def __init__(self, handle: int):
        self._h = C.c_void_p(handle)

    def _apply(self, fn, rows, cols, blocks):
        r = np.ascontiguousarray(rows, np.uint32).reshape(-1)
        c = np.ascontiguousarray(cols, np.uint32).reshape(-1)
        b = np.ascontiguousarray(blocks, np.float32).reshape(-1)
        if len(r) != len(c) or len(b) != 9 * len(r):
            raise ValueError("rows, cols and blocks disagree in length")
        check(fn(self._h, len(r), _ptr(r, u32p), _ptr(c, u32p), _ptr(b, f32p)), fn.__name__)

    def add(self, rows, cols, blocks):
        self._apply(_lib.lib().b3g_matrix_add, rows, cols, blocks)

    def sub(self, rows, cols, blocks):
        self._apply(_lib.lib().b3g_matrix_sub, rows, cols, blocks)

    def set_row(self, row, cols, blocks):
        """Row ``row`` becomes ``cols`` / ``blocks`` head to tail, values as given (``b3g_matrix_set_row``)."""
        c = np.ascontiguousarray(cols, np.uint32).reshape(-1)
        b = np.ascontiguousarray(blocks, np.float32).reshape(-1)
        if len(b) != 9 * len(c):
            raise ValueError("cols and blocks disagree in length")
        check(_lib.lib().b3g_matrix_set_row(self._h, int(row), len(c), _ptr(c, u32p), _ptr(b, f32p)),
              "b3g_matrix_set_row")


@dataclass
class SparseForceSolverData:
    """``b3SparseForceSolverData`` (sparse_force_solver.h:31-39): x and v are read-only (n, 3) views of the
    iterate, f a writable (n, 3) view that starts at zero."""

    h: float
    inv_h: float
    x: np.ndarray
    v: np.ndarray
    f: np.ndarray
    dfdx: SparseMat33
    dfdv: SparseMat33


class SparseForceModel:
    """``b3SparseForceModel``: implement :meth:`ApplyForces`. Any callable taking the data works too."""

    def ApplyForces(self, data: SparseForceSolverData) -> None:  # noqa: N802 (reference name)
        raise NotImplementedError

    def __call__(self, data: SparseForceSolverData) -> None:
        self.ApplyForces(data)


@dataclass
class SolveBEInput:
    """``b3SolveBEInput`` (sparse_force_solver.h:52-81) with its constructor's defaults; M and S are (n, 3, 3)
    or (n, 9) column-major diagonal blocks, z may be None (zero)."""

    h: float = 0.0
    inv_h: float = 0.0
    forceModel: object = None   # noqa: N815 (reference name)
    x0: np.ndarray | None = None
    v0: np.ndarray | None = None
    fe: np.ndarray | None = None
    M: np.ndarray | None = None
    y: np.ndarray | None = None
    S: np.ndarray | None = None
    z: np.ndarray | None = None
    maxIterations: int = 1   # noqa: N815
    tolerance: float = FLT_EPSILON
    maxSubIterations: int = 20   # noqa: N815
    subTolerance: float = FLT_EPSILON   # noqa: N815


@dataclass
class SolveBEOutput:
    """``b3SolveBEOutput`` (sparse_force_solver.h:84-93); the sub-iteration bounds are folded into, as the
    reference folds them (callers start them at 0xFFFFFFFF and 0)."""

    x: np.ndarray | None = None
    v: np.ndarray | None = None
    iterations: int = 0
    error: float = 0.0
    minSubIterations: int = 0xFFFFFFFF   # noqa: N815
    maxSubIterations: int = 0   # noqa: N815


class BESolver:
    """A ``b3g_be`` context on one device: device buffers and the symbolic pattern persist between solves."""

    def __init__(self, device: int = 0, reduction: int = _lib.REDUCE_FP64):
        self._L = _lib.lib()
        self._h = C.c_void_p()
        check(self._L.b3g_be_create(C.byref(self._h), device), "b3g_be_create")
        self.set_reduction(reduction)

    def set_reduction(self, mode: int):
        check(self._L.b3g_be_set_reduction(self._h, mode), "b3g_be_set_reduction")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._L.b3g_be_destroy(h)
            self._h = C.c_void_p()

    def solve(self, output: SolveBEOutput, inp: SolveBEInput) -> SolveBEOutput:
        """``b3SparseSolveBE(&output, &input)``. A force-model exception ends the solve and is raised again."""
        f32 = np.float32
        n = len(inp.x0)
        arr = {k: np.ascontiguousarray(getattr(inp, k), f32).reshape(-1) for k in ("x0", "v0", "fe", "y")}
        M = np.ascontiguousarray(inp.M, f32).reshape(-1)
        S = np.ascontiguousarray(inp.S, f32).reshape(-1)
        z = None if inp.z is None else np.ascontiguousarray(inp.z, f32).reshape(-1)
        if any(len(a) != 3 * n for a in arr.values()) or len(M) != 9 * n or len(S) != 9 * n or (
                z is not None and len(z) != 3 * n):
            raise ValueError("input arrays disagree with n")
        xo = np.zeros(3 * n, f32)
        vo = np.zeros(3 * n, f32)
        model = inp.forceModel
        raised = []

        def cb(user, dp):
            try:
                d = dp.contents
                m = d.n
                data = SparseForceSolverData(
                    h=d.h, inv_h=d.inv_h,
                    x=np.ctypeslib.as_array(d.x, (m, 3)) if m else np.zeros((0, 3), f32),
                    v=np.ctypeslib.as_array(d.v, (m, 3)) if m else np.zeros((0, 3), f32),
                    f=np.ctypeslib.as_array(d.f, (m, 3)) if m else np.zeros((0, 3), f32),
                    dfdx=SparseMat33(d.dfdx), dfdv=SparseMat33(d.dfdv))
                data.x.flags.writeable = False
                data.v.flags.writeable = False
                model(data)
                return _lib.B3G_OK
            except BaseException as exc:   # noqa: BLE001 -- carried across the C boundary, raised below
                raised.append(exc)
                return _lib.B3G_INVALID_ARGUMENT

        fn = _lib.FORCE_MODEL(cb)
        bi = _lib.BEInput()
        bi.h, bi.inv_h = inp.h, inp.inv_h
        bi.force_model = fn
        bi.user = None
        bi.n = n
        bi.x0, bi.v0, bi.fe, bi.y = (_ptr(arr[k], f32p) for k in ("x0", "v0", "fe", "y"))
        bi.M, bi.S = _ptr(M, f32p), _ptr(S, f32p)
        bi.z = _ptr(z, f32p) if z is not None else C.cast(None, f32p)
        bi.max_iterations, bi.tolerance = inp.maxIterations, inp.tolerance
        bi.max_sub_iterations, bi.sub_tolerance = inp.maxSubIterations, inp.subTolerance
        bo = _lib.BEOutput()
        bo.x, bo.v = _ptr(xo, f32p), _ptr(vo, f32p)
        bo.min_sub_iterations = output.minSubIterations
        bo.max_sub_iterations = output.maxSubIterations
        st = self._L.b3g_solve_be(self._h, C.byref(bo), C.byref(bi))
        if raised:
            raise raised[0]
        output.x, output.v = xo.reshape(-1, 3), vo.reshape(-1, 3)
        output.iterations, output.error = bo.iterations, bo.error
        output.minSubIterations, output.maxSubIterations = bo.min_sub_iterations, bo.max_sub_iterations
        if st != _lib.B3G_OK:
            raise B3gError(st, "b3g_solve_be", self._L.b3g_last_error().decode())
        return output


def SparseSolveBE(output: SolveBEOutput, inp: SolveBEInput, solver: BESolver | None = None) -> SolveBEOutput:  # noqa: N802
    """``b3SparseSolveBE(output, input)`` on the GPU (a fresh :class:`BESolver` on device 0 unless given)."""
    return (solver or BESolver()).solve(output, inp)
