"""GPU parity of the broad phase (d2_broad_phase) and the cluster separation
(d2_clusters_build) against the oracle (oracle/d2o_broad.hpp): identical pair
lists and cluster numbering, roll-back states within 1e-12, and the whole
device step (broad phase -> clusters -> step run) against the oracle's
per-cluster steps and contacts."""
import numpy as np
import pytest

import d2o
import delta2_amd as d2
from broad_cases import CUBE_F, CUBE_V, plane_mesh, random_scene, reference_cases
from delta2_amd import pcg_seed, scenes, tree_build

pytestmark = pytest.mark.gpu


def _ctx_for(meshes, records, last=None, groups=None):
    trees = []
    rng = pcg_seed(42, 54)
    for V, F in meshes:
        trees.append(tree_build(V, F, rng))
    c = d2.Context(0, d2.D2_FP64)
    c.upload_trees(trees)
    c.upload_bodies(np.arange(len(meshes), dtype=np.int32), records)
    if last is not None:
        c.upload_last_states(last)
    if groups is not None:
        c.upload_groups(groups)
    return c


def _oracle_bounds(meshes):
    b = [d2o.mesh_bounds(V) for V, _ in meshes]
    return np.array([x[0] for x in b]), np.array([x[1] for x in b])


def test_tree_bounds_equal_mesh_bounds():
    rng = np.random.default_rng(3)
    meshes = [(CUBE_V, CUBE_F), plane_mesh(10), scenes.icosphere(2, 1.3), scenes.ellipsoid_hull(rng)]
    c = _ctx_for(meshes, np.stack([scenes.body_record([1, 0, 0, 0], [0, 0, 0], [0, 0, 0], np.eye(3))] * 4))
    box, rad = c.tree_bounds(len(meshes))
    wb, wr = _oracle_bounds(meshes)
    assert np.array_equal(box, wb)
    assert np.array_equal(rad, wr)
    c.close()


@pytest.mark.parametrize("case", range(6))
def test_reference_broad_phase_counts(case):
    # collision_detection/test_broad_phase.cpp:6-110
    name, s, t, want = reference_cases(d2o.mesh_bounds)[case]
    meshes = {"stationary": [(CUBE_V, CUBE_F)] * 10, "moving": [(CUBE_V, CUBE_F)] * 2,
              "static": [plane_mesh(10), (CUBE_V, CUBE_F)],
              "static2": [plane_mesh(10), (CUBE_V, CUBE_F), plane_mesh(10)]}[name]
    print(f"[broad] reference case {name} t={t}", flush=True)
    c = _ctx_for(meshes, s.records)
    n = c.broad_phase(mode=0, t=t)
    assert n == want
    pairs, mt = c.download_pairs()
    assert np.array_equal(pairs, d2o.broad_phase(s.records, s.aabb6, s.radius, mode=0, t=t))
    assert np.array_equal(mt, np.minimum(s.records[pairs[:, 0], 39], s.records[pairs[:, 1], 39]))
    c.close()


CASES = [
    dict(seed=21, n=400, mode=0, t=0.05),
    dict(seed=22, n=400, mode=1),
    dict(seed=23, n=1500, mode=1, box=20.0, n_static=5),
    dict(seed=24, n=1200, mode=1, n_groups=8),
    dict(seed=25, n=600, mode=1, uneven_times=True),
    dict(seed=26, n=5000, mode=1, box=30.0, n_static=2),  # static sweeps over two segments
    dict(seed=27, n=300, mode=0, t=0.3, n_static=0, vel_sigma=6.0),
]


@pytest.mark.parametrize("k", range(len(CASES)))
def test_broad_phase_matches_oracle(k):
    p = dict(CASES[k])
    mode, t = p.pop("mode"), p.pop("t", 0.0)
    print(f"[broad] case {k} {CASES[k]}", flush=True)
    meshes, rec, last, groups = random_scene(**p)
    c = _ctx_for(meshes, rec, last, groups)
    n = c.broad_phase(mode=mode, t=t)
    got, mt = c.download_pairs()
    box, rad = _oracle_bounds(meshes)
    want = d2o.broad_phase(rec, box, rad, mode=mode, t=t, last=last, groups=groups)
    print(f"[broad] pairs {n} (oracle {want.shape[0]})", flush=True)
    assert want.shape[0] > 0
    assert np.array_equal(got, want)
    assert np.array_equal(mt, np.minimum(rec[got[:, 0], 39], rec[got[:, 1], 39]))
    # the same context again gives the same list
    assert c.broad_phase(mode=mode, t=t) == n
    assert np.array_equal(c.download_pairs()[0], want)
    c.close()


def test_broad_phase_long_partner_lists():
    # one large body among many small ones: its list exceeds what one wave
    # sorts in LDS; a static plate at index 0 sweeps over two segments
    rng = np.random.default_rng(31)
    n = 4200
    V0, F0 = scenes.icosphere(1, 1.0)
    meshes, rec = [plane_mesh(40.0)], [scenes.body_record([1, 0, 0, 0], [10, 10, 2], [0, 0, 0], np.eye(3), static=True)]
    meshes.append((V0 * 9.0, F0))
    rec.append(scenes.body_record([1, 0, 0, 0], [10, 10, 10], [0, 0, 0], np.eye(3)))
    for _ in range(n - 2):
        meshes.append((V0 * 0.3, F0))
        rec.append(scenes.body_record(scenes.random_rotation(rng), rng.uniform(0, 20, 3), rng.normal(0, 1, 3), np.eye(3)))
    rec = np.array(rec)
    rec[:, 14:28] = rec[:, 0:14]
    c = _ctx_for(meshes, rec)
    c.broad_phase(mode=1)
    got, _ = c.download_pairs()
    box, rad = _oracle_bounds(meshes)
    want = d2o.broad_phase(rec, box, rad, mode=1)
    assert np.sum(want[:, 0] == 1) > 512
    assert np.array_equal(got, want)
    c.close()


def test_broad_phase_pair_capacity_grows():
    meshes, rec, last, groups = random_scene(41, n=3000, box=9.0, n_static=0)
    c = _ctx_for(meshes, rec, last)
    n = c.broad_phase(mode=1)  # initial capacity: 8 pairs per body
    box, rad = _oracle_bounds(meshes)
    want = d2o.broad_phase(rec, box, rad, mode=1, last=last)
    print(f"[broad] dense: {n} pairs for {rec.shape[0]} bodies", flush=True)
    assert n > 8 * rec.shape[0]
    assert np.array_equal(c.download_pairs()[0], want)
    c.close()


def test_interpolation_error_is_reported():
    # a body whose last state lies after the other body's time: State::interpolate throws
    meshes, rec, last, _ = random_scene(51, n=200, uneven_times=True)
    box, rad = _oracle_bounds(meshes)
    pairs = d2o.broad_phase(rec, box, rad, mode=1, last=last)
    a, b = pairs[0]
    late = b if rec[b, 13] >= rec[a, 13] else a
    last[late, 13] = rec[late, 13] - 1e-9
    with pytest.raises(RuntimeError):
        d2o.broad_phase(rec, box, rad, mode=1, last=last)
    c = _ctx_for(meshes, rec, last)
    with pytest.raises(d2.D2Error) as e:
        c.broad_phase(mode=1)
    assert e.value.code == d2.D2_ERR_RUNTIME
    c.close()


@pytest.mark.parametrize("uneven", [False, True])
def test_clusters_match_oracle(uneven):
    meshes, rec, last, _ = random_scene(61 + uneven, n=800, box=16.0, n_static=4, uneven_times=uneven)
    print(f"[clusters] uneven={uneven}", flush=True)
    c = _ctx_for(meshes, rec, last)
    c.broad_phase(mode=1)
    pairs, _ = c.download_pairs()
    k = c.build_clusters()
    got = c.download_clusters()
    want = d2o.clusters(rec, pairs, last=last)
    assert k == want["n"]
    for key in ("of_pair", "of_body", "first"):
        assert np.array_equal(got[key], want[key]), key
    assert np.array_equal(got["min_time"], want["min_time"])
    assert np.array_equal(got["step"], want["step"])
    cur = c.download_bodies()[:, 0:14]
    assert np.allclose(cur, want["current"], rtol=1e-12, atol=1e-14)
    if not uneven:
        assert np.array_equal(cur, rec[:, 0:14])
    c.close()


def test_device_step_matches_oracle():
    """Broad phase, clusters and the step on the device against the oracle's
    broad phase, cluster separation and fineCollisionClustersWithTimeStepSelection."""
    from gpu_cases import dense_scene, oracle_scene, with_future
    s = dense_scene()
    rec = s.records.copy()
    rec[::5, 39] = scenes.DT / 3
    rec[:, 14:28] = with_future(rec, rec[:, 39])[:, 14:28]
    c = d2.Context(0, d2.D2_FP64)
    c.upload_trees(s.trees)
    c.upload_bodies(s.tree_of_body, rec)
    box, rad = c.tree_bounds(len(s.trees))
    n = c.broad_phase(mode=1)
    pairs, mt = c.download_pairs()
    want_pairs = d2o.broad_phase(rec, box[s.tree_of_body], rad[s.tree_of_body], mode=1)
    assert np.array_equal(pairs, want_pairs)
    k = c.build_clusters()
    cl = d2o.clusters(rec, pairs)
    assert k == cl["n"]
    n_contacts, steps = c.step_run(scenes.DT, filter=True, want_steps=True)
    print(f"[device step] pairs={n} clusters={k} contacts={n_contacts}", flush=True)
    osc = oracle_scene(s, rec)
    dt32 = np.float32(scenes.DT)
    want_steps = []
    for j in range(k):
        m = cl["of_pair"] == j
        if m.any():
            want_steps.append(osc.cluster_step(pairs[m], dt32))
        else:  # no interactions: the first particle's last step, clamped
            want_steps.append(max(1e-3, min(rec[cl["first"][j], 39], float(dt32))))
    assert np.allclose(steps, want_steps, rtol=1e-12, atol=0)
    body_step = np.array([steps[cl["of_body"][i]] if cl["of_body"][i] >= 0 else 0.0 for i in range(rec.shape[0])])
    got_bodies = c.download_bodies()
    dyn = rec[:, 38] == 0
    want_bodies = with_future(rec, body_step)
    assert np.allclose(got_bodies[dyn, 14:28], want_bodies[dyn, 14:28], rtol=0, atol=1e-14)
    gpu = c.result(with_pairs=True)
    o_off, o_ints, o_reals = oracle_scene(s, got_bodies).narrow(pairs, filter=True)
    assert np.array_equal(gpu.offsets, o_off)
    assert np.array_equal(gpu.ints, o_ints)
    _, _, _, o_pp = oracle_scene(s, rec).ccd(pairs, mt)
    assert np.allclose(gpu.per_pair, o_pp, rtol=1e-9, atol=1e-12)
    c.close()
