"""The many-body workloads at full size, through the C ABI: a batch instance
steps exactly as the same body stepped alone.

* Contact batch: 256 x [128^2-vertex cloth + 16^3-cell tet block] over one
  128^3 SDF, every vertex a sphere fixture (5.45 M particles, 44.1 M blocks),
  captured steps. After the 40 steps that settle the bodies into contact,
  three instances equal standalone copies bit for bit, with the same
  statistics, and contacts are active.
* Uneven batch: 4096 tet boxes (12.6 M particles, 148.8 M blocks). After three
  steps, the first, middle and last instances equal standalone copies bit for
  bit, with the same PCG iteration counts.
"""

import numpy as np
import pytest

from bounce_softbody_amd import _lib, configs
from bounce_softbody_amd.body import Body

pytestmark = pytest.mark.gpu
SUB, SUB_TOL = 50, 1e-5


@pytest.fixture(scope="module", autouse=True)
def _device():
    if _lib.device_count() < 1:
        pytest.fail("no GPU visible to libb3g")


def check_instances(batch, singles, picks, when):
    xb, vb = batch.state()
    xs, vs = batch.split(xb), batch.split(vb)
    stats = batch.instance_stats()
    for s, k in zip(singles, picks):
        x1, v1 = s.state()
        assert np.array_equal(xs[k], x1) and np.array_equal(vs[k], v1), f"instance {k}, {when}"
        st, cc = s.stats()
        assert np.array_equal(stats[k][0], st) and stats[k][1] == cc, f"instance {k} statistics, {when}"


def test_contact_batch_full_size_instances_equal_standalone():
    arrs = configs.contact_batch(256)
    picks = [0, 127, 255]
    batch = Body(arrs)
    singles = [Body(arrs[k]) for k in picks]
    for b in [batch] + singles:
        b.set_graph(True)
    for step in range(40):
        for b in [batch] + singles:
            b.Step(configs.DT, 1, SUB, 0.0, SUB_TOL)
        if step % 10 == 9:
            check_instances(batch, singles, picks, f"step {step}")
    assert all(s.stats()[1] > 0 for s in singles)   # the compared bodies are in contact


def test_batch4096_full_size_instances_equal_standalone():
    arrs = configs.batch_boxes(4096)
    picks = [0, 2047, 4095]
    batch = Body(arrs)
    batch.set_graph(True)
    singles = [Body(arrs[k]) for k in picks]
    for step in range(3):
        for b in [batch] + singles:
            b.Step(configs.DT, 1, SUB, 0.0, SUB_TOL)
        check_instances(batch, singles, picks, f"step {step}")
