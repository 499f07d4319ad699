"""Batched form: several bodies in one handle, stepped by one launch sequence.

Each instance keeps its own Newton / PCG state, convergence and contact list,
and every reduction stays within an instance over the same partition it has
alone, so an instance of a batch reproduces its standalone run bit for bit.
Standalone runs are pinned against the reference elsewhere (sequential mode
against its goldens, FP64 mode per step within 1e-4), so these checks carry
the batch to the same bar.
"""

import numpy as np
import pytest

from bounce_softbody_amd import _lib, configs
from bounce_softbody_amd.body import Body
from tests.golden.make_golden import external_loads

pytestmark = pytest.mark.gpu

SCENES = ["tet_grid", "node_types", "drop_on_box", "cloth_sphere", "cloth_shapes", "cloth_element", "springs",
          "mouse", "sdf", "sdf_inverted", "external", "grid_contact", "mesh_terrain"]


@pytest.fixture(scope="module", autouse=True)
def _device():
    if _lib.device_count() < 1:
        pytest.fail("no GPU visible to libb3g")


def uneven_batch():
    """The parity scenes plus tet grids of different sizes (uneven instances)."""
    arrs = [configs.PARITY_SCENES[n]().arrays() for n in SCENES]
    arrs += [configs.tet_grid(k).arrays() for k in (2, 5, 3)]
    return arrs


def run_both(arrs, mode, steps, it=2, sub=20, stol=1e-5):
    singles = [Body(a) for a in arrs]
    batch = Body(arrs)
    for b in singles + [batch]:
        b.set_reduction(mode)
    ext = [i for i, n in enumerate(SCENES) if n == "external"]
    for step in range(steps):
        if step % 3 == 0:
            loads = []
            for k, a in enumerate(arrs):
                n = len(a["p_type"])
                f, t = external_loads(n) if k in ext else (np.zeros((n, 3), np.float32), np.zeros((n, 3), np.float32))
                if k in ext:
                    singles[k].apply_external(f, t)
                loads.append((f, t))
            batch.apply_external(np.concatenate([l[0] for l in loads]), np.concatenate([l[1] for l in loads]))
        for s in singles:
            s.Step(configs.DT, it, sub, 0.0, stol)
        batch.Step(configs.DT, it, sub, 0.0, stol)
        xb, vb = batch.state()
        for k, (xs, vs) in enumerate(zip(batch.split(xb), batch.split(vb))):
            x1, v1 = singles[k].state()
            assert np.array_equal(xs, x1) and np.array_equal(vs, v1), f"instance {k} differs at step {step}"
    return singles, batch


@pytest.mark.parametrize("mode", [_lib.REDUCE_FP64, _lib.REDUCE_SEQUENTIAL], ids=["fp64", "sequential"])
def test_batch_instances_equal_standalone_bodies(mode):
    arrs = uneven_batch()
    singles, batch = run_both(arrs, mode, steps=10)
    stats = batch.instance_stats()
    assert len(stats) == len(arrs)
    for k, s in enumerate(singles):
        st, cc = s.stats()
        assert np.array_equal(stats[k][0], st) and stats[k][1] == cc, f"instance {k} statistics"
    assert batch.energies() == [s.GetEnergy() for s in singles]
    assert np.array_equal(batch.tet_quats(), np.concatenate([s.tet_quats() for s in singles]))
    # contacts: instance after instance, indices offset into the concatenated creation order
    bp, bw, bf = batch.contacts()
    exp_p, exp_w, exp_f = [], [], []
    p_off = w_off = 0
    for s, a in zip(singles, arrs):
        p, w, f = s.contacts()
        exp_p.append(p + p_off)
        exp_w.append(w + w_off)
        exp_f.append(f)
        p_off += len(a["p_type"])
        w_off += len(a["w_type"])
    assert np.array_equal(bp, np.concatenate(exp_p)) and np.array_equal(bw, np.concatenate(exp_w))
    assert np.array_equal(bf, np.concatenate(exp_f))
    assert len(bp) > 0


def test_batch_convergence_is_per_instance():
    """Instances that need different PCG iteration counts each stop on their own."""
    arrs = [configs.tet_grid(4).arrays(), configs.tet_grid(4, young=1e7).arrays(), configs.PARITY_SCENES["springs"]().arrays()]
    singles, batch = run_both(arrs, _lib.REDUCE_FP64, steps=4, it=1, sub=60, stol=1e-6)
    its = [int(st[2]) for st, _ in batch.instance_stats()]
    assert len(set(its)) > 1, its


def test_batch_checkpoint_replays_bit_exactly():
    arrs = [configs.PARITY_SCENES[n]().arrays() for n in ("drop_on_box", "cloth_sphere", "tet_grid")]
    b = Body(arrs)
    for _ in range(5):
        b.Step(configs.DT, 1, 20, 0.0, 1e-5)
    blob = b.save_state()
    for _ in range(4):
        b.Step(configs.DT, 1, 20, 0.0, 1e-5)
    x1, v1 = b.state()
    c1 = b.contacts()
    b.load_state(blob)
    for _ in range(4):
        b.Step(configs.DT, 1, 20, 0.0, 1e-5)
    x2, v2 = b.state()
    assert np.array_equal(x1, x2) and np.array_equal(v1, v2)
    assert all(np.array_equal(p, q) for p, q in zip(c1, b.contacts()))
    single = Body(arrs[0])
    with pytest.raises(_lib.B3gError):
        single.load_state(blob)


def test_batch_reports_the_failing_instance():
    good = configs.PARITY_SCENES["springs"]().arrays()
    bad = dict(good)
    bad["sp_p"] = np.array(good["sp_p"], np.uint32).copy()
    bad["sp_p"][0] = len(good["p_type"]) + 5
    with pytest.raises(_lib.B3gError, match="instance 1"):
        Body([good, bad])


@pytest.mark.parametrize("other", [_lib.REDUCE_FP64_FUSED, _lib.REDUCE_FP64], ids=["fused", "default"])
def test_fused_reductions_equal_separate_finalize(other):
    """B3G_REDUCE_FP64_FUSED forms the same sums in the same order inside the
    consuming kernels, and the default mode picks one of the two forms:
    trajectories equal the separate-kernel form bit for bit."""
    arrs = [configs.PARITY_SCENES[n]().arrays() for n in ("tet_grid", "cloth_sphere", "springs", "sdf")]
    a, b = Body(arrs), Body(arrs)
    a.set_reduction(_lib.REDUCE_FP64_SEPARATE)
    b.set_reduction(other)
    for _ in range(8):
        a.Step(configs.DT, 2, 30, 0.0, 1e-6)
        b.Step(configs.DT, 2, 30, 0.0, 1e-6)
        xa, va = a.state()
        xb, vb = b.state()
        assert np.array_equal(xa, xb) and np.array_equal(va, vb)
    assert [s[0].tolist() for s in a.instance_stats()] == [s[0].tolist() for s in b.instance_stats()]


def test_contact_workload_batch_equals_oracle():
    """The contact workload at small size (3 bodies of cloth + tet block over one shared SDF grid): in
    sequential mode every instance reproduces the oracle bit for bit, contacts and normal forces included,
    and the FP64-reduced batch equals standalone bodies bit for bit."""
    import oracle
    arrs = configs.contact_batch(3, cloth_cells=11, tet_cells=3, sdf_res=24)
    assert len({a["w_voxels"][0].ctypes.data for a in arrs}) == 1   # one grid, uploaded once
    batch = Body(arrs)
    batch.set_reduction(_lib.REDUCE_SEQUENTIAL)
    refs = [oracle.OracleBody(a) for a in arrs]
    touching = 0
    for step in range(15):
        batch.Step(configs.DT, 1, 30, 0.0, 1e-5)
        for o in refs:
            o.step(configs.DT, 1, 30, 0.0, 1e-5)
        xb, vb = batch.state()
        for k, (xs, vs) in enumerate(zip(batch.split(xb), batch.split(vb))):
            xo, vo = refs[k].state()
            assert np.array_equal(xs, xo) and np.array_equal(vs, vo), f"instance {k} differs at step {step}"
    bp, bw, bf = batch.contacts()
    p_off = w_off = 0
    exp = [[], [], []]
    for o, a in zip(refs, arrs):
        p, w, f = o.contacts()
        exp[0].append(p + p_off)
        exp[1].append(w + w_off)
        exp[2].append(f)
        touching += int(np.count_nonzero(f))
        p_off += len(a["p_type"])
        w_off += len(a["w_type"])
    assert np.array_equal(bp, np.concatenate(exp[0])) and np.array_equal(bw, np.concatenate(exp[1]))
    assert np.array_equal(bf, np.concatenate(exp[2]))
    assert touching > 0
    run_both(arrs, _lib.REDUCE_FP64, steps=6, it=1, sub=30, stol=1e-5)


def test_batch_sharing_one_mesh():
    """Instances passing the same mesh arrays share one device copy of the mesh and its tree; every instance still
    equals its standalone body bit for bit (sequential mode: the reference's own trajectory)."""
    a = configs.PARITY_SCENES["grid_contact"]().arrays()
    b = configs.PARITY_SCENES["mesh_terrain"]().arrays()
    arrs = [a, b, a, b, a]
    assert arrs[0]["w_mesh_vertices"][0] is arrs[2]["w_mesh_vertices"][0]
    run_both(arrs, _lib.REDUCE_SEQUENTIAL, steps=8, it=2, sub=20)
