"""Host-side checks of libb3g: exported symbols and the symbolic pattern (no GPU needed)."""

import os
import re

import numpy as np
import pytest

from bounce_softbody_amd import _lib, configs
import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "b3g.h")).read()
    return sorted(set(re.findall(r"^(?:b3g_status|const char\*)\s+(b3g_\w+)\s*\(", src, re.M)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared_functions()
    assert len(names) >= 20
    for name in names:
        assert hasattr(L, name), name
    assert set(names) == set(_lib.EXPORTS)


@pytest.mark.parametrize("name", sorted(configs.PARITY_SCENES))
def test_pattern_is_reference_row_order(name):
    """pA's blocks, row by row, in the order the reference's linked lists hold them."""
    arr = configs.PARITY_SCENES[name]().arrays()
    rowptr, cols = _lib.pattern(arr)
    orow, ocol, *_ = oracle.OracleBody(arr).assemble(configs.DT)
    assert np.array_equal(rowptr, orow) and np.array_equal(cols, ocol)


def test_pattern_block_count_tet_grid_10():
    rowptr, cols = _lib.pattern(configs.tet_grid(10).arrays())
    assert len(rowptr) == 1332 and len(cols) == 15191


def test_invalid_descriptions_are_rejected():
    arr = configs.tet_grid(2).arrays()
    bad = dict(arr)
    bad["tet_p"] = arr["tet_p"].copy()
    bad["tet_p"][0] = 10 ** 6
    with pytest.raises(_lib.B3gError, match="out of range"):
        _lib.pattern(bad)
    bad = dict(arr)
    bad["f_index"] = arr["f_index"].copy()
    bad["f_index"][1] = bad["f_index"][0]
    with pytest.raises(_lib.B3gError):
        _lib.pattern(bad)
    bad = dict(arr)
    bad["p_type"] = arr["p_type"].copy()
    bad["p_type"][0] = 7
    with pytest.raises(_lib.B3gError):
        _lib.pattern(bad)


def test_invalid_meshes_are_rejected():
    """Mesh world fixtures: vertex indices (wings may be 0xFFFFFFFF) must be in range, the scale non-zero
    (mesh_shape.cpp:94-96), the mesh non-empty."""
    from bounce_softbody_amd.scene import SceneBuilder, STATIC
    def scene(tris, scale=(1.0, 1.0, 1.0)):
        b = SceneBuilder()
        b.add_particles(np.zeros((1, 3), np.float32), types=np.full(1, STATIC, np.uint8))
        b.add_world_mesh(np.array([[0, 0, 0], [1, 0, 0], [0, 0, 1]], np.float32), tris, scale=scale)
        return b.arrays()
    _lib.pattern(scene(np.array([[0, 1, 2, 0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF]], np.uint32)))
    with pytest.raises(_lib.B3gError, match="out of range"):
        _lib.pattern(scene(np.array([[0, 1, 3]], np.uint32)))
    with pytest.raises(_lib.B3gError, match="out of range"):
        _lib.pattern(scene(np.array([[0, 1, 2, 5, 0xFFFFFFFF, 0xFFFFFFFF]], np.uint32)))
    with pytest.raises(_lib.B3gError, match="scale"):
        _lib.pattern(scene(np.array([[0, 1, 2]], np.uint32), scale=(1.0, 0.0, 1.0)))
    with pytest.raises(_lib.B3gError, match="triangles"):
        _lib.pattern(scene(np.zeros((0, 3), np.uint32)))
