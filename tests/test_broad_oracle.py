"""CPU: the broad-phase and cluster oracle (oracle/d2o_broad.hpp) pinned by
the reference's own broad-phase tests, closed forms and properties that do
not restate its formulas."""
import numpy as np
import pytest
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components

import d2o
from broad_cases import CUBE_V, plane_mesh, random_scene, reference_cases


def test_mesh_bounds_match_reference_fixture():
    # model/test_particle.cpp:17-19: the cube's AABB is [-1, 1]^3, sphere bound sqrt(3)
    box, r = d2o.mesh_bounds(CUBE_V)
    assert np.array_equal(box, np.array([-1, -1, -1, 1, 1, 1], dtype=np.float32))
    assert r == np.sqrt(3.0)
    box, r = d2o.mesh_bounds(plane_mesh(10)[0])
    assert np.array_equal(box, np.array([-5, -5, 0, 5, 5, 0], dtype=np.float32))
    assert r == np.sqrt(50.0)


@pytest.mark.parametrize("case", range(6))
def test_reference_broad_phase_counts(case):
    # collision_detection/test_broad_phase.cpp:6-110 (counts 9; 0, 1; 0, 1, 2)
    name, s, t, want = reference_cases(d2o.mesh_bounds)[case]
    p = d2o.broad_phase(s.records, s.aabb6, s.radius, mode=0, t=t)
    assert p.shape[0] == want, (name, t, p)
    assert np.all(p[:, 0] < p[:, 1])


def test_reference_empty_scene():
    # test_broad_phase.cpp:113-120
    p = d2o.broad_phase(np.zeros((0, 40)), np.zeros((0, 6)), np.zeros(0), mode=0, t=1.0)
    assert p.shape == (0, 2)


def _bounds(meshes):
    b = [d2o.mesh_bounds(V) for V, _ in meshes]
    return np.array([x[0] for x in b]), np.array([x[1] for x in b])


def test_resting_bodies_pair_by_sphere_overlap():
    # all velocities zero: a pair is reported exactly when the bounding spheres
    # (radius + geo_eps) overlap (sphereSphereFirstIntersect returns 0 then and -1 otherwise)
    meshes, rec, last, _ = random_scene(11, n=200, n_static=0)
    rec[:, 7:10] = 0.0
    rec[:, 18:21] = rec[:, 4:7]
    box, rad = _bounds(meshes)
    p = d2o.broad_phase(rec, box, rad, mode=0, t=1.0 / 60)
    R = rad + rec[:, 37]
    d = np.linalg.norm(rec[:, None, 4:7] - rec[None, :, 4:7], axis=2)
    i, j = np.nonzero(np.triu(d * d - (R[:, None] + R[None, :]) ** 2 < 0.0, 1))
    want = np.stack([i, j], 1)
    assert np.array_equal(p, want)


def test_fixed_step_pairs_grow_with_the_step():
    # a longer step sweeps larger boxes and admits later first contacts
    meshes, rec, last, _ = random_scene(12, n=250)
    box, rad = _bounds(meshes)
    prev = set()
    for t in (0.0, 0.01, 0.05, 0.2, 0.5):
        cur = set(map(tuple, d2o.broad_phase(rec, box, rad, mode=0, t=t)))
        assert prev <= cur
        prev = cur
    assert len(prev) > 0


def test_pairs_pass_both_stages():
    # every reported pair has overlapping boxes and a first contact in [0, t)
    meshes, rec, last, _ = random_scene(13, n=250)
    box, rad = _bounds(meshes)
    t = 0.2
    p = d2o.broad_phase(rec, box, rad, mode=0, t=t)
    assert p.shape[0] > 10
    for a, b in p:
        assert not (rec[a, 38] and rec[b, 38])
        s = d2o.sphere_sphere(rec[a, 4:7], rec[a, 7:10], rad[a] + rec[a, 37], rec[b, 4:7], rec[b, 7:10],
                              rad[b] + rec[b, 37])
        assert 0.0 <= s < t


def test_groups_split_the_scene():
    meshes, rec, last, groups = random_scene(14, n=300, n_groups=4)
    box, rad = _bounds(meshes)
    got = d2o.broad_phase(rec, box, rad, mode=1, last=last, groups=groups)
    want = []
    for g in range(4):
        idx = np.nonzero(groups == g)[0]
        sub = d2o.broad_phase(rec[idx], box[idx], rad[idx], mode=1, last=last[idx])
        want += [(idx[a], idx[b]) for a, b in sub]
    assert sorted(map(tuple, got)) == sorted(want)


def test_individual_time_with_equal_times_is_the_sphere_rule():
    # equal current times: no interpolation, the window is max(last steps)
    meshes, rec, last, _ = random_scene(15, n=250)
    box, rad = _bounds(meshes)
    p = d2o.broad_phase(rec, box, rad, mode=1, last=last)
    for a, b in p:
        s = d2o.sphere_sphere(rec[a, 4:7], rec[a, 7:10], rad[a] + rec[a, 37], rec[b, 4:7], rec[b, 7:10],
                              rad[b] + rec[b, 37])
        assert 0.0 <= s < max(rec[a, 39], rec[b, 39])


def test_interpolate_closed_forms():
    cur = np.zeros(14)
    last = np.zeros(14)
    th = 0.8  # rotation about z by th between last (identity) and cur
    last[0] = 1.0
    cur[0:4] = [np.cos(th / 2), 0.0, 0.0, np.sin(th / 2)]
    cur[4:7] = [2.0, 4.0, -6.0]
    cur[7:10] = [1.0, 0.0, 0.0]
    cur[13], last[13] = 2.0, 1.0
    for tt in (1.0, 1.25, 1.5, 2.0):
        o = d2o.interpolate(cur, last, tt)
        f = tt - 1.0
        assert o[13] == tt
        assert np.allclose(o[4:7], f * cur[4:7], rtol=0, atol=1e-15)
        assert np.allclose(o[0:4], [np.cos(f * th / 2), 0.0, 0.0, np.sin(f * th / 2)], rtol=0, atol=1e-15)
    # the shorter arc: a quaternion and its negation are the same rotation
    neg = cur.copy()
    neg[0:4] = -cur[0:4]
    o = d2o.interpolate(neg, last, 1.5)
    assert np.allclose(o[0:4], [np.cos(th / 4), 0.0, 0.0, np.sin(th / 4)], atol=1e-15)
    with pytest.raises(RuntimeError):
        d2o.interpolate(cur, last, 0.5)
    with pytest.raises(RuntimeError):
        d2o.interpolate(cur, last, 2.5)
    assert np.array_equal(d2o.interpolate(cur, cur, 2.0), cur)


def _scipy_clusters(rec, pairs):
    n = rec.shape[0]
    st = rec[:, 38] != 0
    keep = ~st[pairs[:, 0]] & ~st[pairs[:, 1]]
    a, b = pairs[keep, 0], pairs[keep, 1]
    _, lab = connected_components(coo_matrix((np.ones(a.size), (a, b)), shape=(n, n)), directed=False)
    # clusters: components with a non-static member, ordered by their smallest member
    first = {}
    for i in range(n):
        if not st[i] and lab[i] not in first:
            first[lab[i]] = i
    order = sorted(first, key=lambda k: first[k])
    cid = {k: c for c, k in enumerate(order)}
    of_body = np.array([-1 if st[i] else cid[lab[i]] for i in range(n)])
    return of_body, [first[k] for k in order]


def test_clusters_match_connected_components():
    meshes, rec, last, _ = random_scene(16, n=400, box=14.0, n_static=4)
    box, rad = _bounds(meshes)
    pairs = d2o.broad_phase(rec, box, rad, mode=1, last=last)
    C = d2o.clusters(rec, pairs, last=last)
    of_body, first = _scipy_clusters(rec, pairs)
    assert C["n"] == len(first)
    assert np.array_equal(C["of_body"], of_body)
    assert np.array_equal(C["first"], first)
    dyn = np.where(rec[pairs[:, 0], 38] != 0, pairs[:, 1], pairs[:, 0])
    assert np.array_equal(C["of_pair"], of_body[dyn])
    for c in range(C["n"]):
        m = of_body == c
        assert C["step"][c] == rec[m, 39].min()
        assert C["min_time"][c] == rec[m, 13].min()
    # equal times: the roll-back changes nothing
    assert np.array_equal(C["current"], rec[:, 0:14])


def test_cluster_rollback_uses_the_cluster_minimum():
    meshes, rec, last, _ = random_scene(17, n=300, uneven_times=True)
    box, rad = _bounds(meshes)
    pairs = d2o.broad_phase(rec, box, rad, mode=1, last=last)
    C = d2o.clusters(rec, pairs, last=last)
    for i in range(rec.shape[0]):
        if rec[i, 38]:
            assert np.array_equal(C["current"][i], rec[i, 0:14])
            continue
        want = d2o.interpolate(rec[i, 0:14], last[i], C["min_time"][C["of_body"][i]])
        assert np.array_equal(C["current"][i], want)
    # the earliest member of every cluster keeps its state (its rotation up to
    # the quaternion's sign: slerp takes the shorter arc)
    for c in range(C["n"]):
        m = np.nonzero(C["of_body"] == c)[0]
        k = m[np.argmin(rec[m, 13])]
        got = C["current"][k]
        assert np.allclose(got[4:14], rec[k, 4:14], rtol=0, atol=1e-15)
        q = rec[k, 0:4]
        assert np.allclose(got[0:4], q, atol=1e-15) or np.allclose(got[0:4], -q, atol=1e-15)
