"""Sanitizer builds for tests/test_host_sanitizers.py (test infrastructure, not the product).

* tests/native/_build/libb3g_hostasan.so: libb3g's sources with AddressSanitizer and UndefinedBehaviorSanitizer
  on the host side only (-Xarch_host; the device code is compiled as usual and never launched by the tests).
  Host-only entry points (the description checks and the symbolic pattern behind b3g_pattern, b3g_sdf_grid) run
  under the sanitizers on a machine without a GPU.
* oracle/_asan/liboracle_asan.so: the oracle's restatement with the same sanitizers (make -C oracle asan).
"""

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from bounce_softbody_amd import build as hipbuild  # noqa: E402

OUT = os.path.join(HERE, "_build", "libb3g_hostasan.so")
ORACLE_ASAN = os.path.join(ROOT, "oracle", "_asan", "liboracle_asan.so")


def build_host_asan() -> str:
    srcs = [os.path.join(hipbuild.CSRC, s) for s in hipbuild.SOURCES]
    flags = [f for f in hipbuild.FLAGS if f != "-O3"]
    # -Xarch_host before each sanitizer: host code only
    san = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
           "-Xarch_host", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer", "-g", "-O1"]
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    cmd = [hipbuild.hipcc(), *flags, *san, "-I", os.path.join(ROOT, "include"), "-I",
           hipbuild.CSRC, *srcs, "-o", OUT + ".tmp"]
    subprocess.run(cmd, check=True, stderr=subprocess.DEVNULL)
    os.replace(OUT + ".tmp", OUT)
    return OUT


def build_oracle_asan() -> str:
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "asan"], check=True)
    return ORACLE_ASAN


def clang_asan_runtime() -> str:
    import glob
    c = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    if not c:
        raise FileNotFoundError("clang's AddressSanitizer runtime")
    return c[-1]


def gcc_asan_runtime() -> str:
    return subprocess.check_output([shutil.which("gcc") or "gcc", "-print-file-name=libasan.so"], text=True).strip()


if __name__ == "__main__":
    print(build_host_asan())
    print(build_oracle_asan())
