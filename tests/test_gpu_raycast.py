"""b3Body::RayCastSingle on the device (b3g_ray_cast) against the reference's goldens and the oracle.

The oracle's brute-force RayCastSingle is pinned bit for bit to the reference
build (goldens in tests/golden/ray_cast.npz, written by its tree query), so the
device results are checked to the same bar: fixture index, fraction and normal
equal bit for bit. Ties (equal fractions) go to the lowest creation index on
both sides; no golden ray has a tie.
"""

import os

import numpy as np
import pytest

from bounce_softbody_amd import _lib, configs
from bounce_softbody_amd.body import Body
import oracle
from tests.golden.make_golden import RAY_SCENES, rays_for

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "ray_cast.npz")


@pytest.fixture(scope="module", autouse=True)
def _device():
    if _lib.device_count() < 1:
        pytest.fail("no GPU visible to libb3g")


def same(a, b):
    return all(np.array_equal(u, w) for u, w in zip(a, b))


@pytest.mark.parametrize("name", RAY_SCENES)
def test_ray_cast_matches_reference_golden(name):
    g = np.load(GOLDEN)
    body = Body(configs.PARITY_SCENES[name]().arrays())
    body.set_state(g[f"{name}/x"], g[f"{name}/v"])
    idx, frac, nrm = body.RayCastSingle(g[f"{name}/p1"], g[f"{name}/p2"])
    assert np.array_equal(idx, g[f"{name}/index"])
    assert np.array_equal(frac, g[f"{name}/fraction"]) and np.array_equal(nrm, g[f"{name}/normal"])


def test_ray_cast_follows_the_steps():
    """Casts between steps see the current positions: after every step the device agrees with the oracle at the
    device's state (the tree is rebuilt only when particles moved)."""
    arr = configs.cloth_on_sphere(48, size=1.6, height=0.5).arrays()
    body = Body(arr)
    o = oracle.OracleBody(arr)
    for step in range(4):
        x, v = body.state()
        p1, p2 = rays_for(x, seed=300 + step, n=600)
        o.set_state(x, v)
        a = body.RayCastSingle(p1, p2)
        assert same(a, o.ray_cast(p1, p2)), f"step {step}"
        assert same(a, body.RayCastSingle(p1, p2))   # a second query on the same state (tree reused)
        assert (a[0] >= 0).sum() > 150
        body.Step(configs.DT, 1, 30, 0.0, 1e-5)


@pytest.mark.parametrize("cells", [255, 511], ids=["cloth256", "cloth512"])
def test_ray_cast_large_cloth(cells):
    """A 256^2-vertex cloth (130,050 triangle fixtures) and the benchmark's 512^2 cloth (522,242) after a few
    steps: 800 rays equal the oracle's brute force."""
    arr = configs.cloth_on_sphere(cells).arrays()
    body = Body(arr)
    for _ in range(5):
        body.Step(configs.DT, 1, 50, 0.0, 1e-5)
    x, v = body.state()
    o = oracle.OracleBody(arr)
    o.set_state(x, v)
    p1, p2 = rays_for(x, seed=77, n=800)
    a = body.RayCastSingle(p1, p2)
    assert same(a, o.ray_cast(p1, p2))
    assert (a[0] >= 0).sum() > 300


def test_ray_cast_batch_instances():
    """Each instance of a batch answers like the standalone body, with indices into the concatenated triangle-
    fixture arrays; an instance without triangle fixtures reports no hit."""
    arrs = [configs.PARITY_SCENES[n]().arrays() for n in ("cloth_sphere", "springs", "tet_grid", "drop_on_box")]
    batch = Body(arrs)
    singles = [Body(a) for a in arrs]
    for _ in range(3):
        batch.Step(configs.DT, 1, 20, 0.0, 1e-5)
        for s in singles:
            s.Step(configs.DT, 1, 20, 0.0, 1e-5)
    off = 0
    for k, (s, a) in enumerate(zip(singles, arrs)):
        x, _ = s.state()
        p1, p2 = rays_for(x, seed=500 + k, n=300)
        i1, f1, n1 = s.RayCastSingle(p1, p2)
        ib, fb, nb = batch.RayCastSingle(p1, p2, instance=k)
        assert np.array_equal(ib, np.where(i1 >= 0, i1 + off, -1)), f"instance {k}"
        assert np.array_equal(fb, f1) and np.array_equal(nb, n1)
        if len(a["tf_density"]) == 0:
            assert (ib < 0).all()
        off += len(a["tf_density"])


def test_ray_cast_validates():
    body = Body(configs.PARITY_SCENES["cloth_sphere"]().arrays())
    with pytest.raises(_lib.B3gError):
        body.RayCastSingle(np.zeros((1, 3)), np.ones((1, 3)), instance=1)
    idx, frac, nrm = body.RayCastSingle(np.zeros((0, 3)), np.zeros((0, 3)))
    assert len(idx) == 0


def _device_cast(body, p1, p2, instance=0):
    n = len(p1)
    d1 = _lib.DeviceArray.of(np.asarray(p1, np.float32))
    d2 = _lib.DeviceArray.of(np.asarray(p2, np.float32))
    di, df, dn = _lib.DeviceArray(8 * n), _lib.DeviceArray(4 * n), _lib.DeviceArray(12 * n)
    body.ray_cast_device(n, d1.address, d2.address, di.address, df.address, dn.address, instance)
    return di.download(np.int64, n), df.download(np.float32, n), dn.download(np.float32, (n, 3))


def test_ray_cast_device_arrays_equal_host_arrays():
    """b3g_ray_cast_device (rays and results in device memory) answers exactly as b3g_ray_cast, in a single body
    and in a batch instance."""
    arr = configs.cloth_on_sphere(64, size=1.6, height=0.5).arrays()
    body = Body(arr)
    for _ in range(3):
        body.Step(configs.DT, 1, 30, 0.0, 1e-5)
    x, _ = body.state()
    p1, p2 = rays_for(x, seed=901, n=3000)
    assert same(_device_cast(body, p1, p2), body.RayCastSingle(p1, p2))
    arrs = [configs.PARITY_SCENES[n]().arrays() for n in ("springs", "cloth_sphere")]
    batch = Body(arrs)
    batch.Step(configs.DT, 1, 20, 0.0, 1e-5)
    x, _ = batch.state()
    p1, p2 = rays_for(x, seed=902, n=500)
    assert same(_device_cast(batch, p1, p2, instance=1), batch.RayCastSingle(p1, p2, instance=1))


def test_ray_cast_device_rejects_host_arrays():
    """Host arrays passed to the device entry point are refused before anything runs on the GPU."""
    body = Body(configs.PARITY_SCENES["cloth_sphere"]().arrays())
    p = np.zeros((4, 3), np.float32)
    idx = np.zeros(4, np.int64)
    fr = np.zeros(4, np.float32)
    with pytest.raises(_lib.B3gError, match="not device memory"):
        body.ray_cast_device(4, p.ctypes.data, p.ctypes.data, idx.ctypes.data, fr.ctypes.data)
    idx2, frac2, _ = body.RayCastSingle(p, p + 1.0)   # the body still answers
    assert len(idx2) == 4
