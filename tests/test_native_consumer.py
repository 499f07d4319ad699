"""A compiled C++ caller of include/b3g.h (tests/native/b3g_consumer.cpp, built by build() against libb3g.so):
it builds a body from a description, steps it on a schedule and writes the state after each step, checking every
status outside any assert. On the GPU its states equal the reference's fixtures (tests/golden/, written by the
reference built from its own sources) bit for bit, step by step, in the reference's summation order.
"""

import os
import struct
import subprocess

import numpy as np
import pytest

from bounce_softbody_amd import _lib, configs
from bounce_softbody_amd.build import CONSUMER
from tests.golden.make_golden import digest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
DESC_ORDER = [("p_type", np.uint8), ("p_pos", np.float32), ("p_vel", np.float32), ("p_damping", np.float32),
              ("s_particle", np.uint32), ("s_radius", np.float32), ("s_friction", np.float32),
              ("tf_p", np.uint32), ("tf_v", np.float32), ("tf_density", np.float32), ("tf_radius", np.float32),
              ("tf_friction", np.float32),
              ("ttf_p", np.uint32), ("ttf_v", np.float32), ("ttf_density", np.float32), ("ttf_radius", np.float32),
              ("ttf_friction", np.float32),
              ("f_type", np.uint32), ("f_index", np.uint32),
              ("tet_p", np.uint32), ("tet_v", np.float32), ("tet_param", np.float32),
              ("tri_p", np.uint32), ("tri_v", np.float32), ("tri_param", np.float32),
              ("st_p", np.uint32), ("st_param", np.float32), ("sh_p", np.uint32), ("sh_param", np.float32),
              ("sp_p", np.uint32), ("sp_param", np.float32), ("mo_p", np.uint32), ("mo_param", np.float32),
              ("w_type", np.uint32), ("w_param", np.float32), ("w_friction", np.float32)]


def write_scene(path, arr, schedule):
    with open(path, "wb") as f:
        for key, dt in DESC_ORDER:
            a = np.ascontiguousarray(arr[key], dt).reshape(-1)
            f.write(struct.pack("<IQ", a.itemsize, a.size))
            f.write(a.tobytes())
        f.write(np.asarray(arr["gravity"], np.float32).tobytes())
        f.write(struct.pack("<Q", len(schedule)))
        for dt, it, sub, tol, stol in schedule:
            f.write(struct.pack("<fIIff", dt, int(it), int(sub), tol, stol))


def test_consumer_is_built_and_checks_its_input(tmp_path):
    """No GPU needed: usage and a truncated scene file are refused before any device call."""
    assert os.path.exists(CONSUMER), "tests/native/_build/b3g_consumer is built by __graft_entry__.build()"
    assert subprocess.run([CONSUMER], capture_output=True, timeout=60).returncode == 64
    bad = tmp_path / "bad.bin"
    bad.write_bytes(b"\x01\x00")
    r = subprocess.run([CONSUMER, str(bad), str(tmp_path / "out.bin")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 65 and "malformed" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["tet_grid", "cloth_sphere", "drop_on_box", "springs", "mouse", "cloth_element"])
def test_consumer_reproduces_reference_golden(tmp_path, name):
    if _lib.device_count() < 1:
        pytest.fail("no GPU visible to libb3g")
    g = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    assert not g["external"].any()
    arr = configs.PARITY_SCENES[name]().arrays()
    scene, out = tmp_path / "scene.bin", tmp_path / "out.bin"
    write_scene(scene, arr, [tuple(r) for r in g["schedule"]])
    r = subprocess.run([CONSUMER, str(scene), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    n = len(arr["p_type"])
    states = np.fromfile(out, np.float32).reshape(len(g["schedule"]), 2, n, 3)
    for k, (x, v) in enumerate(states):
        assert np.array_equal(digest(x, v), g["digests"][k]), f"{name}: step {k}"
    assert np.array_equal(states[-1, 0], g["x"]) and np.array_equal(states[-1, 1], g["v"])
