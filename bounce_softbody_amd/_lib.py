"""ctypes binding of libb3g.so (include/b3g.h).

The library is built in-tree by :mod:`bounce_softbody_amd.build`. Loading
fails loudly when it is missing: there is no CPU fallback.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libb3g.so")

u32p = C.POINTER(C.c_uint32)
f32p = C.POINTER(C.c_float)
u8p = C.POINTER(C.c_uint8)

B3G_OK = 0
B3G_INVALID_ARGUMENT, B3G_HIP_ERROR, B3G_OUT_OF_MEMORY, B3G_DEGENERATE_ELEMENT, B3G_UNSUPPORTED, B3G_NOT_COMMITTED = (
    1, 2, 3, 4, 5, 6)
B3G_NON_POSITIVE_DIAGONAL, B3G_NON_FINITE = 7, 8
FAULT_DIAGONAL, FAULT_NON_FINITE = 1, 2
REDUCE_FP64, REDUCE_SEQUENTIAL, REDUCE_FP64_FUSED, REDUCE_FP64_SEPARATE = 0, 1, 2, 3
SOLVER_PCG, SOLVER_REFINED_FP32, SOLVER_REFINED_BF16 = 0, 1, 2
STATUS = {0: "ok", 1: "invalid argument", 2: "HIP error", 3: "out of device memory", 4: "degenerate element",
          5: "unsupported", 6: "body not built", 7: "non-positive Jacobi diagonal", 8: "non-finite state"}


class B3gError(RuntimeError):
    def __init__(self, status: int, where: str, detail: str):
        super().__init__(f"{where}: {STATUS.get(status, status)}: {detail}")
        self.status = status


class BodyDesc(C.Structure):
    """``b3g_body_desc`` (include/b3g.h)."""

    _fields_ = [
        ("n_particles", C.c_uint32), ("p_type", u8p), ("p_pos", f32p), ("p_vel", f32p), ("p_damping", f32p),
        ("n_spheres", C.c_uint32), ("s_particle", u32p), ("s_radius", f32p), ("s_friction", f32p),
        ("n_trifix", C.c_uint32), ("tf_p", u32p), ("tf_v", f32p), ("tf_density", f32p), ("tf_radius", f32p),
        ("tf_friction", f32p),
        ("n_tetfix", C.c_uint32), ("ttf_p", u32p), ("ttf_v", f32p), ("ttf_density", f32p), ("ttf_radius", f32p),
        ("ttf_friction", f32p),
        ("n_forces", C.c_uint32), ("f_type", u32p), ("f_index", u32p),
        ("n_tet", C.c_uint32), ("tet_p", u32p), ("tet_v", f32p), ("tet_param", f32p),
        ("n_tri", C.c_uint32), ("tri_p", u32p), ("tri_v", f32p), ("tri_param", f32p),
        ("n_stretch", C.c_uint32), ("st_p", u32p), ("st_param", f32p),
        ("n_shear", C.c_uint32), ("sh_p", u32p), ("sh_param", f32p),
        ("n_spring", C.c_uint32), ("sp_p", u32p), ("sp_param", f32p),
        ("n_mouse", C.c_uint32), ("mo_p", u32p), ("mo_param", f32p),
        ("n_world", C.c_uint32), ("w_type", u32p), ("w_param", f32p), ("w_friction", f32p),
        ("w_voxels", C.POINTER(f32p)),
        ("gravity", C.c_float * 3),
        ("w_mesh_vertices", C.POINTER(f32p)), ("w_mesh_triangles", C.POINTER(u32p)), ("w_mesh_counts", u32p),
    ]


class StepStats(C.Structure):
    _fields_ = [("newton_iterations", C.c_uint32), ("min_sub_iterations", C.c_uint32),
                ("max_sub_iterations", C.c_uint32), ("contact_count", C.c_uint32),
                ("refinement_sweeps", C.c_uint32), ("relative_residual", C.c_float), ("faults", C.c_uint32)]


def _ptr(a, t):
    return a.ctypes.data_as(t) if a is not None and a.size else C.cast(None, t)


def make_desc(arr: dict):
    """``b3g_body_desc`` over the arrays of :meth:`SceneBuilder.arrays`; returns (desc, keepalive)."""
    keep = {k: (np.ascontiguousarray(v) if isinstance(v, np.ndarray) else v) for k, v in arr.items()}
    d = BodyDesc()
    d.n_particles = len(keep["p_type"])
    d.p_type, d.p_pos, d.p_vel, d.p_damping = (_ptr(keep["p_type"], u8p), _ptr(keep["p_pos"], f32p),
                                               _ptr(keep["p_vel"], f32p), _ptr(keep["p_damping"], f32p))
    d.n_spheres = len(keep["s_particle"])
    d.s_particle, d.s_radius, d.s_friction = (_ptr(keep["s_particle"], u32p), _ptr(keep["s_radius"], f32p),
                                              _ptr(keep["s_friction"], f32p))
    d.n_trifix = len(keep["tf_density"])
    for k, t in (("tf_p", u32p), ("tf_v", f32p), ("tf_density", f32p), ("tf_radius", f32p), ("tf_friction", f32p)):
        setattr(d, k, _ptr(keep[k], t))
    d.n_tetfix = len(keep["ttf_density"])
    for k, t in (("ttf_p", u32p), ("ttf_v", f32p), ("ttf_density", f32p), ("ttf_radius", f32p), ("ttf_friction", f32p)):
        setattr(d, k, _ptr(keep[k], t))
    d.n_forces = len(keep["f_type"])
    d.f_type, d.f_index = _ptr(keep["f_type"], u32p), _ptr(keep["f_index"], u32p)
    d.n_tet = len(keep["tet_param"]) // 3
    d.tet_p, d.tet_v, d.tet_param = _ptr(keep["tet_p"], u32p), _ptr(keep["tet_v"], f32p), _ptr(keep["tet_param"], f32p)
    d.n_tri = len(keep["tri_param"]) // 6
    d.tri_p, d.tri_v, d.tri_param = _ptr(keep["tri_p"], u32p), _ptr(keep["tri_v"], f32p), _ptr(keep["tri_param"], f32p)
    d.n_stretch = len(keep["st_param"]) // 12
    d.st_p, d.st_param = _ptr(keep["st_p"], u32p), _ptr(keep["st_param"], f32p)
    d.n_shear = len(keep["sh_param"]) // 8
    d.sh_p, d.sh_param = _ptr(keep["sh_p"], u32p), _ptr(keep["sh_param"], f32p)
    d.n_spring = len(keep["sp_param"]) // 3
    d.sp_p, d.sp_param = _ptr(keep["sp_p"], u32p), _ptr(keep["sp_param"], f32p)
    d.n_mouse = len(keep["mo_param"]) // 6
    d.mo_p, d.mo_param = _ptr(keep["mo_p"], u32p), _ptr(keep["mo_param"], f32p)
    d.n_world = len(keep["w_type"])
    d.w_type, d.w_param, d.w_friction = (_ptr(keep["w_type"], u32p), _ptr(keep["w_param"], f32p),
                                         _ptr(keep["w_friction"], f32p))
    vox = keep["w_voxels"]
    arrp = (f32p * max(1, len(vox)))(*[(_ptr(v, f32p) if v is not None else C.cast(None, f32p)) for v in vox])
    keep["_vox"] = arrp
    d.w_voxels = C.cast(arrp, C.POINTER(f32p))
    d.gravity = (C.c_float * 3)(*[float(g) for g in keep["gravity"]])
    mv = keep.get("w_mesh_vertices", [None] * len(vox))
    mt = keep.get("w_mesh_triangles", [None] * len(vox))
    mv_arr = (f32p * max(1, len(mv)))(*[(_ptr(v, f32p) if v is not None else C.cast(None, f32p)) for v in mv])
    mt_arr = (u32p * max(1, len(mt)))(*[(_ptr(t, u32p) if t is not None else C.cast(None, u32p)) for t in mt])
    counts = np.ascontiguousarray(keep.get("w_mesh_counts", np.zeros(2 * len(vox), np.uint32)), np.uint32)
    keep["_mv"], keep["_mt"], keep["_mc"] = mv_arr, mt_arr, counts
    d.w_mesh_vertices = C.cast(mv_arr, C.POINTER(f32p))
    d.w_mesh_triangles = C.cast(mt_arr, C.POINTER(u32p))
    d.w_mesh_counts = _ptr(counts, u32p)
    return d, keep


_lib = None

EXPORTS = [
    "b3g_status_string", "b3g_last_error", "b3g_device_count", "b3g_body_create", "b3g_body_destroy",
    "b3g_body_build", "b3g_body_build_batch", "b3g_get_instance_count", "b3g_step", "b3g_get_state", "b3g_set_state", "b3g_apply_external", "b3g_get_masses",
    "b3g_get_stats", "b3g_get_energy", "b3g_get_contacts", "b3g_get_tet_quats", "b3g_assemble", "b3g_get_system",
    "b3g_synchronize", "b3g_time_spmv", "b3g_set_solver", "b3g_set_graph", "b3g_get_system_size", "b3g_pattern",
    "b3g_solve_cg", "b3g_spmv", "b3g_set_reduction", "b3g_save_state", "b3g_load_state", "b3g_ray_cast", "b3g_ray_cast_device",
    "b3g_sdf_grid", "b3g_build_sdf", "b3g_body_rebuild", "b3g_set_action_forces", "b3g_get_action_forces",
    "b3g_solve_cg_static", "b3g_time_spmv_cold", "b3g_set_early_exit", "b3g_get_schedule", "b3g_set_schedule",
    "b3g_autotune", "b3g_matrix_add", "b3g_matrix_sub", "b3g_matrix_set_row", "b3g_be_create", "b3g_be_destroy", "b3g_be_set_reduction",
    "b3g_be_input_defaults", "b3g_solve_be",
]


class Schedule(C.Structure):
    """``b3g_schedule`` (include/b3g.h): the PCG's SpMV tile shape, vector row-block size and FP64 sum place."""

    _fields_ = [("tile_blocks", C.c_uint32), ("tile_rows", C.c_uint32), ("row_block_rows", C.c_uint32),
                ("fused", C.c_int32), ("max_groups", C.c_uint32), ("persist_tiles", C.c_uint32),
                ("persist_groups", C.c_uint32)]

    def as_tuple(self):
        return tuple(getattr(self, f) for f in SCHEDULE_FIELDS)

    def __repr__(self):
        return "Schedule(%s)" % ", ".join("%s=%d" % (f, getattr(self, f)) for f in SCHEDULE_FIELDS)


SCHEDULE_FIELDS = ("tile_blocks", "tile_rows", "row_block_rows", "fused", "max_groups", "persist_tiles",
                   "persist_groups")
DEFAULT_SCHEDULE = (1024, 256, 1024, -1, 0, 0, 0)


class ForceData(C.Structure):
    """``b3g_force_data`` (include/b3g.h): b3SparseForceSolverData."""

    _fields_ = [("h", C.c_float), ("inv_h", C.c_float), ("n", C.c_uint32), ("x", f32p), ("v", f32p), ("f", f32p),
                ("dfdx", C.c_void_p), ("dfdv", C.c_void_p)]


# b3g_force_model: b3g_status (*)(void* user, const b3g_force_data* data)
FORCE_MODEL = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(ForceData))


class BEInput(C.Structure):
    """``b3g_be_input`` (include/b3g.h): b3SolveBEInput."""

    _fields_ = [("h", C.c_float), ("inv_h", C.c_float), ("force_model", FORCE_MODEL), ("user", C.c_void_p),
                ("n", C.c_uint32), ("x0", f32p), ("v0", f32p), ("fe", f32p), ("M", f32p), ("y", f32p), ("S", f32p),
                ("z", f32p), ("max_iterations", C.c_uint32), ("tolerance", C.c_float),
                ("max_sub_iterations", C.c_uint32), ("sub_tolerance", C.c_float)]


class BEOutput(C.Structure):
    """``b3g_be_output`` (include/b3g.h): b3SolveBEOutput."""

    _fields_ = [("x", f32p), ("v", f32p), ("iterations", C.c_uint32), ("error", C.c_float),
                ("min_sub_iterations", C.c_uint32), ("max_sub_iterations", C.c_uint32)]


class Carry(C.Structure):
    """``b3g_carry`` (include/b3g.h)."""

    _fields_ = [("particle_prev", C.POINTER(C.c_int64)), ("tet_prev", C.POINTER(C.c_int64)),
                ("sphere_prev", C.POINTER(C.c_int64)), ("world_prev", C.POINTER(C.c_int64))]


class DeviceArray:
    """A device buffer from the HIP runtime libb3g.so is linked against (hipMalloc / hipMemcpy / hipFree), for
    the entry points that take device pointers. Another HIP runtime in the same process (a framework's bundled
    copy) keeps its own allocations, which this library's runtime does not recognise."""

    _hip = None

    @classmethod
    def runtime(cls):
        if cls._hip is None:
            lib()   # libb3g.so first: the runtime it bound is the one returned for its soname
            h = C.CDLL("libamdhip64.so.7")
            h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            h.hipFree.argtypes = [C.c_void_p]
            h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            cls._hip = h
        return cls._hip

    def __init__(self, nbytes: int):
        self.nbytes = max(1, int(nbytes))
        self.ptr = C.c_void_p()
        rc = self.runtime().hipMalloc(C.byref(self.ptr), self.nbytes)
        if rc != 0:
            raise B3gError(rc, "hipMalloc", f"hipMalloc of {self.nbytes} bytes failed")

    @classmethod
    def of(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        rc = self.runtime().hipMemcpy(self.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes, 1)   # host to device
        if rc != 0:
            raise B3gError(rc, "hipMemcpy", "host to device copy failed")

    def download(self, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype)
        assert out.nbytes <= self.nbytes
        rc = self.runtime().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, out.nbytes, 2)   # device to host
        if rc != 0:
            raise B3gError(rc, "hipMemcpy", "device to host copy failed")
        return out

    @property
    def address(self) -> int:
        return self.ptr.value

    def free(self):
        if self.ptr and self._hip is not None:
            self._hip.hipFree(self.ptr)
        self.ptr = C.c_void_p()

    def __del__(self):
        # at interpreter exit the runtime may already be torn down: leave the memory to the process's end
        import sys
        if not sys.is_finalizing():
            self.free()


def lib():
    """Loads libb3g.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(f"{LIB_PATH} is missing: build it with `python -m bounce_softbody_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    sig = {
        "b3g_status_string": (C.c_char_p, [C.c_int]),
        "b3g_last_error": (C.c_char_p, []),
        "b3g_device_count": (C.c_int, [C.POINTER(C.c_int)]),
        "b3g_body_create": (C.c_int, [C.POINTER(vp), C.c_int]),
        "b3g_body_destroy": (C.c_int, [vp]),
        "b3g_body_build": (C.c_int, [vp, C.POINTER(BodyDesc)]),
        "b3g_body_build_batch": (C.c_int, [vp, C.POINTER(BodyDesc), C.c_uint32]),
        "b3g_get_instance_count": (C.c_int, [vp, u32p]),
        "b3g_ray_cast": (C.c_int, [vp, C.c_uint32, C.c_uint32, f32p, f32p, C.POINTER(C.c_int64), f32p, f32p]),
        "b3g_ray_cast_device": (C.c_int, [vp, C.c_uint32, C.c_uint32, vp, vp, vp, vp, vp]),
        "b3g_sdf_grid": (C.c_int, [f32p, C.c_uint32, f32p, C.c_float, f32p, u32p]),
        "b3g_body_rebuild": (C.c_int, [vp, C.c_void_p, C.c_uint32, C.POINTER(Carry)]),
        "b3g_set_action_forces": (C.c_int, [vp, C.c_int]),
        "b3g_get_action_forces": (C.c_int, [vp, C.c_uint32, f32p, C.c_uint64]),
        "b3g_build_sdf": (C.c_int, [C.c_int, f32p, C.c_uint32, u32p, C.c_uint32, C.c_uint32, f32p, C.c_float, f32p,
                                    C.c_uint64]),
        "b3g_step": (C.c_int, [vp, C.c_float, C.c_uint32, C.c_uint32, C.c_float, C.c_float]),
        "b3g_get_state": (C.c_int, [vp, f32p, f32p]),
        "b3g_set_state": (C.c_int, [vp, f32p, f32p]),
        "b3g_apply_external": (C.c_int, [vp, f32p, f32p]),
        "b3g_get_masses": (C.c_int, [vp, f32p]),
        "b3g_get_stats": (C.c_int, [vp, C.POINTER(StepStats)]),
        "b3g_get_energy": (C.c_int, [vp, f32p]),
        "b3g_get_contacts": (C.c_int, [vp, u32p, u32p, u32p, f32p, C.c_uint32]),
        "b3g_get_tet_quats": (C.c_int, [vp, f32p]),
        "b3g_assemble": (C.c_int, [vp, C.c_float, u32p]),
        "b3g_get_system": (C.c_int, [vp, u32p, u32p, f32p, f32p, f32p]),
        "b3g_synchronize": (C.c_int, [vp]),
        "b3g_time_spmv": (C.c_int, [vp, C.c_uint32, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "b3g_time_spmv_cold": (C.c_int, [vp, C.c_uint32, C.c_uint64, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "b3g_set_early_exit": (C.c_int, [vp, C.c_int]),
        "b3g_get_schedule": (C.c_int, [vp, C.POINTER(Schedule)]),
        "b3g_set_schedule": (C.c_int, [vp, C.POINTER(Schedule)]),
        "b3g_autotune": (C.c_int, [vp, C.c_char_p, C.POINTER(Schedule), C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "b3g_set_solver": (C.c_int, [vp, C.c_int, C.c_uint32, C.c_float]),
        "b3g_set_graph": (C.c_int, [vp, C.c_int]),
        "b3g_get_system_size": (C.c_int, [vp, u32p, u32p]),
        "b3g_pattern": (C.c_int, [C.POINTER(BodyDesc), u32p, u32p, u32p, C.c_uint32]),
        "b3g_solve_cg": (C.c_int, [C.c_int, C.c_uint32, u32p, u32p, f32p, f32p, f32p, C.c_uint32, C.c_float, u32p]),
        "b3g_spmv": (C.c_int, [C.c_int, C.c_uint32, u32p, u32p, f32p, f32p, f32p]),
        "b3g_solve_cg_static": (C.c_int, [C.c_int, vp, C.c_uint32, C.c_uint32, vp, vp, vp, vp, vp, C.c_uint32, C.c_float,
                                          vp]),
        "b3g_set_reduction": (C.c_int, [vp, C.c_int]),
        "b3g_matrix_add": (C.c_int, [vp, C.c_uint32, u32p, u32p, f32p]),
        "b3g_matrix_sub": (C.c_int, [vp, C.c_uint32, u32p, u32p, f32p]),
        "b3g_matrix_set_row": (C.c_int, [vp, C.c_uint32, C.c_uint32, u32p, f32p]),
        "b3g_be_create": (C.c_int, [C.POINTER(vp), C.c_int]),
        "b3g_be_destroy": (C.c_int, [vp]),
        "b3g_be_set_reduction": (C.c_int, [vp, C.c_int]),
        "b3g_be_input_defaults": (C.c_int, [C.POINTER(BEInput)]),
        "b3g_solve_be": (C.c_int, [vp, C.POINTER(BEOutput), C.POINTER(BEInput)]),
        "b3g_save_state": (C.c_int, [vp, C.c_void_p, C.POINTER(C.c_uint64)]),
        "b3g_load_state": (C.c_int, [vp, C.c_void_p, C.c_uint64]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(status: int, where: str):
    if status != B3G_OK:
        raise B3gError(status, where, lib().b3g_last_error().decode())


def device_count() -> int:
    n = C.c_int(0)
    st = lib().b3g_device_count(C.byref(n))
    return n.value if st == B3G_OK else 0


def pattern(arrays: dict):
    """pA's block pattern (rowptr, cols) in the reference's row order, computed on the host."""
    d, keep = make_desc(arrays)
    n = d.n_particles
    nnz = C.c_uint32(0)
    rowptr = np.zeros(n + 1, np.uint32)
    check(lib().b3g_pattern(C.byref(d), C.byref(nnz), _ptr(rowptr, u32p), None, 0), "b3g_pattern")
    cols = np.zeros(max(1, nnz.value), np.uint32)
    check(lib().b3g_pattern(C.byref(d), C.byref(nnz), _ptr(rowptr, u32p), _ptr(cols, u32p), nnz.value), "b3g_pattern")
    return rowptr, cols[: nnz.value]


def solve_cg(rowptr, cols, vals, b, x0, max_iterations, tolerance, device=0):
    """Device PCG (b3SparseSolveCG) on a list-order CSR matrix."""
    n = len(rowptr) - 1
    rowptr = np.ascontiguousarray(rowptr, np.uint32)
    cols = np.ascontiguousarray(cols, np.uint32)
    vals = np.ascontiguousarray(vals, np.float32).reshape(-1)
    b = np.ascontiguousarray(b, np.float32).reshape(-1)
    x = np.ascontiguousarray(x0, np.float32).reshape(-1).copy()
    it = C.c_uint32(0)
    check(lib().b3g_solve_cg(device, n, _ptr(rowptr, u32p), _ptr(cols, u32p), _ptr(vals, f32p), _ptr(b, f32p),
                             _ptr(x, f32p), max_iterations, C.c_float(tolerance), C.byref(it)), "b3g_solve_cg")
    return x.reshape(-1, 3), it.value


CG_STATIC_MAX_ROWS = 65536
CG_STATIC_MAX_ITERATIONS = 4095
CG_STATIC_INVALID = 0xFFFFFFFF


def solve_cg_static(n, nnz, rowptr, cols, planes, b, x, max_iterations, tolerance, iterations=None, device=0,
                    stream=None):
    """b3g_solve_cg_static on device arrays (DeviceArray or raw addresses): the PCG with no allocation, its
    scratch a fixed device array of the library. ``planes``: the blocks as nine planes (entry c of block k at
    c * nnz + k). Asynchronous on ``stream`` (None: the null stream)."""
    addr = lambda a: None if a is None else (a.address if isinstance(a, DeviceArray) else int(a))
    check(lib().b3g_solve_cg_static(device, stream, n, nnz, addr(rowptr), addr(cols), addr(planes), addr(b), addr(x),
                                    max_iterations, C.c_float(tolerance), addr(iterations)), "b3g_solve_cg_static")


def spmv(rowptr, cols, vals, x, device=0):
    n = len(rowptr) - 1
    rowptr = np.ascontiguousarray(rowptr, np.uint32)
    cols = np.ascontiguousarray(cols, np.uint32)
    vals = np.ascontiguousarray(vals, np.float32).reshape(-1)
    x = np.ascontiguousarray(x, np.float32).reshape(-1)
    y = np.zeros(3 * n, np.float32)
    check(lib().b3g_spmv(device, n, _ptr(rowptr, u32p), _ptr(cols, u32p), _ptr(vals, f32p), _ptr(x, f32p),
                         _ptr(y, f32p)), "b3g_spmv")
    return y.reshape(-1, 3)


def build_sdf(vertices, triangles, cell, extension: float = 1.0, device: int = 0):
    """``b3BuildSDF(sdf, mesh, cellSize, aabbVolumeExtension)`` on the device: (box lower, box upper, (W, H, D),
    voxels as an array of shape (D, H, W)). ``triangles`` is (m, 3) or (m, 6) (b3Triangle with wings)."""
    v = np.ascontiguousarray(vertices, np.float32).reshape(-1)
    t = np.ascontiguousarray(triangles, np.uint32)
    stride = t.shape[1]
    c = np.ascontiguousarray(cell, np.float32).reshape(3)
    box = np.zeros(6, np.float32)
    dims = np.zeros(3, np.uint32)
    check(lib().b3g_sdf_grid(_ptr(v, f32p), len(v) // 3, _ptr(c, f32p), C.c_float(extension), _ptr(box, f32p),
                             _ptr(dims, u32p)), "b3g_sdf_grid")
    W, H, D = (int(x) for x in dims)
    vox = np.zeros(W * H * D, np.float32)
    tf = t.reshape(-1)
    check(lib().b3g_build_sdf(device, _ptr(v, f32p), len(v) // 3, _ptr(tf, u32p), len(tf) // stride, stride,
                              _ptr(c, f32p), C.c_float(extension), _ptr(vox, f32p), vox.size), "b3g_build_sdf")
    return box[:3].copy(), box[3:].copy(), (W, H, D), vox.reshape(D, H, W)
