"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer, on the CPU.

* libb3g's host-side set-up, built for the host only with both sanitizers (tests/native/build_sanitized.py):
  the description checks and the symbolic pattern of every parity scene and of a box batch (b3g_pattern), the
  rejected descriptions, and b3g_sdf_grid. Nothing launches a kernel.
* The oracle's restatement with both sanitizers: the golden replays and the operator-level backward Euler.

Each runs in a subprocess with the sanitizer runtime preloaded; any report ends it with a non-zero status.
"""

import os
import subprocess
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(HERE, "native"))

import build_sanitized as bs  # noqa: E402

SAN_ENV = {"ASAN_OPTIONS": "detect_leaks=0:halt_on_error=1:abort_on_error=0",
           "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1"}


def stale(out, deps):
    return not os.path.exists(out) or any(os.path.getmtime(d) > os.path.getmtime(out) for d in deps)


def run(code_or_args, preload, extra_env=None, timeout=900):
    env = dict(os.environ, **SAN_ENV, LD_PRELOAD=preload, PYTHONPATH=ROOT + os.pathsep + HERE)
    env.update(extra_env or {})
    args = code_or_args if isinstance(code_or_args, list) else [sys.executable, "-c", code_or_args]
    p = subprocess.run(args, env=env, cwd=ROOT, capture_output=True, text=True, timeout=timeout)
    assert p.returncode == 0, (p.stdout[-4000:], p.stderr[-6000:])
    assert "ERROR: AddressSanitizer" not in p.stderr and "runtime error:" not in p.stderr, p.stderr[-6000:]
    return p.stdout


HOST_DRIVER = r"""
import ctypes as C, sys
import numpy as np
from bounce_softbody_amd import _lib, configs
L = C.CDLL(sys.argv[1])
L.b3g_pattern.argtypes = [C.POINTER(_lib.BodyDesc), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                          C.POINTER(C.c_uint32), C.c_uint32]
L.b3g_sdf_grid.argtypes = [C.POINTER(C.c_float), C.c_uint32, C.POINTER(C.c_float), C.c_float,
                           C.POINTER(C.c_float), C.POINTER(C.c_uint32)]
u32p = C.POINTER(C.c_uint32)
def pattern(arr):
    d, keep = _lib.make_desc(arr)
    nnz = C.c_uint32(0)
    n = len(arr["p_type"])
    rp = np.zeros(n + 1, np.uint32)
    st = L.b3g_pattern(C.byref(d), C.byref(nnz), rp.ctypes.data_as(u32p), None, 0)
    if st != 0:
        return st, None, None
    cols = np.zeros(max(1, nnz.value), np.uint32)
    st = L.b3g_pattern(C.byref(d), C.byref(nnz), rp.ctypes.data_as(u32p), cols.ctypes.data_as(u32p), nnz.value)
    return st, rp, cols[: nnz.value]
out = {}
scenes = {n: f().arrays() for n, f in configs.PARITY_SCENES.items()}
scenes["tet_grid_10"] = configs.tet_grid(10).arrays()
for k, a in enumerate(configs.batch_boxes(8)):
    scenes[f"box{k}"] = a
for name, arr in scenes.items():
    st, rp, cols = pattern(arr)
    assert st == 0, (name, st)
    out[name + "/rowptr"], out[name + "/cols"] = rp, cols
# rejected descriptions: the checks run before anything is built
arr = configs.tet_grid(2).arrays()
for key, fix in (("tet_p", lambda a: a.__setitem__(0, 10 ** 6)), ("p_type", lambda a: a.__setitem__(0, 7))):
    bad = dict(arr)
    bad[key] = arr[key].copy()
    fix(bad[key])
    assert pattern(bad)[0] == 1, key
bad = dict(arr)
bad["f_index"] = arr["f_index"].copy()
bad["f_index"][1] = bad["f_index"][0]
assert pattern(bad)[0] != 0
v = np.random.default_rng(0).normal(size=(500, 3)).astype(np.float32)
box = np.zeros(6, np.float32)
dims = np.zeros(3, np.uint32)
cell = np.array([0.05, 0.04, 0.03], np.float32)
assert L.b3g_sdf_grid(v.ctypes.data_as(C.POINTER(C.c_float)), 500, cell.ctypes.data_as(C.POINTER(C.c_float)),
                      C.c_float(0.5), box.ctypes.data_as(C.POINTER(C.c_float)), dims.ctypes.data_as(u32p)) == 0
out["sdf/box"], out["sdf/dims"] = box, dims
np.savez(sys.argv[2], **out)
print("host sanitizer run ok", len(scenes))
"""


def test_libb3g_host_setup_under_sanitizers(tmp_path):
    deps = [os.path.join(bs.hipbuild.CSRC, s) for s in os.listdir(bs.hipbuild.CSRC)] + [
        os.path.join(ROOT, "include", "b3g.h")]
    lib = bs.OUT if not stale(bs.OUT, deps) else bs.build_host_asan()
    out = str(tmp_path / "patterns.npz")
    run([sys.executable, "-c", HOST_DRIVER, lib, out], bs.clang_asan_runtime())
    # the sanitized build computes what the product computes
    from bounce_softbody_amd import _lib, configs
    g = np.load(out)
    for name, f in configs.PARITY_SCENES.items():
        rp, cols = _lib.pattern(f().arrays())
        assert np.array_equal(rp, g[name + "/rowptr"]) and np.array_equal(cols, g[name + "/cols"]), name


def test_oracle_under_sanitizers():
    src = os.path.join(ROOT, "oracle")
    deps = [os.path.join(src, f) for f in ("oracle_capi.cpp", "cpu_ref.hpp", "cpu_ref_math.hpp", "cpu_ref_sparse.hpp",
                                           "scene.h")]
    lib = bs.ORACLE_ASAN if not stale(bs.ORACLE_ASAN, deps) else bs.build_oracle_asan()
    out = run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
               "tests/test_oracle_golden.py", "tests/test_oracle_be.py", "tests/test_tet_k0.py"],
              bs.gcc_asan_runtime(), {"B3O_ORACLE_SO": lib})
    assert "passed" in out and "failed" not in out, out[-2000:]
