"""Known-answer tests of the CPU oracle, taken from the reference's own tests.

Every vector here is a fixture of the reference test-suite (file:line cited);
margins are the reference's.  Both float and double instantiations are run
where the reference runs both (TEMPLATE_TEST_CASE ... float, double).

Euler-angle fixtures: the reference's expected values for poses built with
``transformationMatrix(euler, pos)`` were produced in the R = Rz*Ry*Rx
convention (test_utils.cpp:68-79 expects Ry*Rx for euler (x, y, 0)), while
the code (basic_utils.h:70-77, Eigen AngleAxis X*Y*Z) yields Rx*Ry*Rz.  The
oracle follows the code; fixtures are rebuilt here in the convention their
expected values were produced with (``fixture_pose``), and a separate test
pins the code's convention against its closed form.
"""
import numpy as np
import pytest

import d2o

PRECS = ["f64", "f32"]
TRI1 = np.array([-1, 0, 0, 1, 0, 0, 0, -2, 0], dtype=float)


def _rx(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[1, 0, 0], [0, c, -s], [0, s, c]])


def _ry(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])


def _rz(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])


def fixture_pose(euler, pos):
    m = np.eye(4)
    m[:3, :3] = _rz(euler[2]) @ _ry(euler[1]) @ _rx(euler[0])
    m[:3, 3] = pos
    return m


def transformed(tri, M, prec):
    t = np.asarray(tri, dtype=float).reshape(3, 3)
    return np.concatenate([d2o.transform_point(v, M, prec) for v in t])


def deg(x):
    return 2.0 * 3.14159265358979323846 * x / 360.0  # basic_utils.h:101-105


# --- common/test_utils.cpp -------------------------------------------------
@pytest.mark.parametrize("prec", PRECS)
def test_transformation_matrix_rotation_z(prec):
    # test_utils.cpp:30-47 (single-axis: convention independent)
    m = d2o.transformation_matrix_euler([0, 0, 3.14159 / 2.0], [0, 0, 0], prec)
    want = np.array([[0, -1, 0, 0], [1, 0, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1.0]])
    assert np.all(np.abs(m - want) <= 1e-3)
    assert np.all(np.abs(d2o.transform_point([1, 0, 0], m, prec) - [0, 1, 0]) < 1e-3)


@pytest.mark.parametrize("prec", PRECS)
def test_transformation_matrix_code_convention(prec):
    # The code's product AngleAxis(x,X)*AngleAxis(y,Y)*AngleAxis(z,Z) is Rx*Ry*Rz.
    e = np.array([0.5235, 0.3491, 0.2])
    m = d2o.transformation_matrix_euler(e, [0, 1, 1.2651], prec)
    want = _rx(e[0]) @ _ry(e[1]) @ _rz(e[2])
    tol = 1e-12 if prec == "f64" else 1e-6
    assert np.allclose(m[:3, :3], want, atol=tol)
    assert np.allclose(m[:3, 3], [0, 1, 1.2651], atol=tol)
    # test_utils.cpp:68-79 expected matrix is the fixture convention Ry*Rx:
    ref_expected = np.array([[0.9396926164627075, 0.1710100769996643, 0.29619812965393066],
                             [0.0, 0.8660253882408142, -0.5],
                             [-0.3420201241970062, 0.46984633803367615, 0.813797652721405]])
    assert np.allclose(fixture_pose([0.5235, 0.3491, 0.0], [0, 0, 0])[:3, :3], ref_expected, atol=1e-3)


# --- common/test_triangle.cpp ----------------------------------------------
@pytest.mark.parametrize("prec", PRECS)
def test_triangle_interaction(prec):
    a, b, c = [0, 0, 0.0], [1, 0, 0.0], [0, 1, 0.0]
    tri = np.array(a + b + c)
    # triangle.h:90-103 barycentrics of the vertices are exact (test_triangle.cpp:57-68)
    assert list(d2o.barycentric(tri, a, prec)) == [1.0, 0.0, 0.0]
    assert list(d2o.barycentric(tri, b, prec)) == [0.0, 1.0, 0.0]
    assert list(d2o.barycentric(tri, c, prec)) == [0.0, 0.0, 1.0]
    d, e = [-1.0, 1, 0], [-1.0, -1, 0]
    tri2 = np.array(b + d + e)
    t, tinf, isect, p = d2o.intersect_segment(tri2, [-1, 1, 1, 1, -1, -1], prec)  # :73-85
    assert isect and not tinf
    assert np.all(np.abs(p) < 1e-3)
    inside, q = d2o.project_to_plane(tri2, [2, 2, 1], prec)  # :87-93
    assert not inside
    assert np.all(np.abs(q - [2, 2, 0]) < 1e-3)
    assert abs(d2o.dist_to_point(tri2, [2, 0, 0], prec) - 1.0) < 1e-5  # :95-103


def test_intersect_segment_regression():
    # test_continuous_comparison.cpp:265-275 (pointTriCCDLinear error case)
    tri = [50, 50, 0, -50, 50, 0, 50, -50, 0]
    t, tinf, isect, _ = d2o.intersect_segment(tri, [2, -1, 0.29, 2, -1, -0.81])
    assert isect


# --- common/test_edge.cpp --------------------------------------------------
@pytest.mark.parametrize("prec", PRECS)
def test_edge_interaction(prec):
    ab = [0, 1, 0, 0, -1, 0.0]
    cd = [1, 0, 1, 1, 0, -1.0]
    ef = [0, 0, 1, 0, 0, 0.0]
    dist = lambda e, f: d2o.edge_closest(e, f, prec)[0]
    assert dist(ab, cd) == pytest.approx(1.0)
    assert dist(ab, ef) == pytest.approx(0.0, abs=1e-6)
    assert dist(ef, cd) == pytest.approx(1.0)
    assert dist(ab, cd) == pytest.approx(dist(cd, ab))
    assert dist(ab, ef) == pytest.approx(dist(ef, ab), abs=1e-6)
    assert dist(ef, cd) == pytest.approx(dist(cd, ef))
    _, p, q = d2o.edge_closest(ab, cd, prec)
    _, ps, qs = d2o.edge_closest(cd, ab, prec)
    assert np.all(np.abs(p - [0, 0, 0]) < 1e-3) and np.all(np.abs(q - [1, 0, 0]) < 1e-3)
    assert np.all(np.abs(p - qs) < 1e-3) and np.all(np.abs(q - ps) < 1e-3)
    _, p, q = d2o.edge_closest(ab, ef, prec)
    _, ps, qs = d2o.edge_closest(ef, ab, prec)
    assert np.all(np.abs(p) < 1e-3) and np.all(np.abs(q) < 1e-3)
    assert np.all(np.abs(p - qs) < 1e-3) and np.all(np.abs(q - ps) < 1e-3)
    _, p, q = d2o.edge_closest(ef, cd, prec)
    _, ps, qs = d2o.edge_closest(cd, ef, prec)
    assert p[0] == pytest.approx(0.0, abs=1e-6) and p[1] == pytest.approx(0.0, abs=1e-6)
    assert 0.0 - 1e-6 <= p[2] <= 1.0 + 1e-6
    assert q[0] == pytest.approx(1.0) and q[1] == pytest.approx(0.0, abs=1e-6)
    assert 0.0 - 1e-6 <= q[2] <= 1.0 + 1e-6
    assert np.all(np.abs(p - qs) < 1e-3) and np.all(np.abs(q - ps) < 1e-3)


# --- collision_detection/test_comparison.cpp -------------------------------
B_POSE = ([-0.1745, 0.0, 0.0873], [0.0, 0.0, 0.0])


@pytest.mark.parametrize("prec", PRECS)
def test_comparison_vert_face(prec):
    # test_comparison.cpp:8-34
    A = transformed(TRI1, fixture_pose([0.5235, 0.3491, 0.0], [0.0, 1.0, 1.2651]), prec)
    B = transformed(TRI1, fixture_pose(*B_POSE), prec)
    d, P, Q = d2o.comparison_dist(A, B, prec)
    Pw, Qw = np.array([-0.342, -0.732, 0.326]), np.array([-0.339, -0.767, 0.129])
    assert np.all(np.abs(P - Pw) < 0.01)
    assert np.all(np.abs(Q - Qw) < 0.01)
    assert abs(d - np.linalg.norm(Pw - Qw)) < 0.02


@pytest.mark.parametrize("prec", PRECS)
def test_comparison_edge_edge(prec):
    # test_comparison.cpp:36-62
    A = transformed(TRI1, fixture_pose([1.323341, 0.301492, 0.887539], [-1.0284, -1.0942, 1.1695]), prec)
    B = transformed(TRI1, fixture_pose(*B_POSE), prec)
    d, P, Q = d2o.comparison_dist(A, B, prec)
    Pw, Qw = np.array([-0.6545, -0.9378, 0.2656]), np.array([-0.5089, -0.8695, 0.1449])
    assert np.all(np.abs(P - Pw) < 0.01)
    assert np.all(np.abs(Q - Qw) < 0.01)
    assert abs(d - np.linalg.norm(Pw - Qw)) < 0.02


@pytest.mark.parametrize("prec", PRECS)
def test_comparison_intersecting(prec):
    # test_comparison.cpp:64-88
    A = transformed(TRI1, fixture_pose([1.323341, 0.301492, 0.887539], [0.052572, -0.153374, 1.0443]), prec)
    B = transformed(TRI1, fixture_pose(*B_POSE), prec)
    d, _, _ = d2o.comparison_dist(A, B, prec)
    assert abs(d) < 1e-4


PYRAMID_V = np.array([[1.0, 1.0, 0.0], [-1.0, 1.0, 0.0], [-1.0, -1.0, 0.0], [1.0, -1.0, 0.0], [0.0, 0.0, 0.1]])
PYRAMID_F = np.array([[0, 1, 4], [1, 2, 4], [2, 3, 4], [3, 0, 4]])
# common/primitive_geo.cpp plane(1.0)
PLANE_V = np.array([[0.5, 0.5, 0.0], [0.5, -0.5, 0.0], [-0.5, 0.5, 0.0], [-0.5, -0.5, 0.0]])
PLANE_F = np.array([[0, 2, 1], [1, 2, 3]])


def body_record(trans=(0, 0, 0), vel=(0, 0, 0), fut_trans=None, geo_eps=0.01, static=False, last_step=1.0):
    r = np.zeros(40)
    r[0] = 1.0
    r[4:7] = trans
    r[7:10] = vel
    r[14] = 1.0
    r[18:21] = trans if fut_trans is None else fut_trans
    r[21:24] = vel
    r[28:37] = np.eye(3).ravel()
    r[37] = geo_eps
    r[38] = 1.0 if static else 0.0
    r[39] = last_step
    return r


def test_bucket_comparison_pyramid_plane():
    # test_comparison.cpp:90-128: pair_used[0] == 8, filtered size 1
    rng = d2o.pcg_seed(42, 54)
    t_a = d2o.build_tree(PYRAMID_V, PYRAMID_F, rng)
    t_b = d2o.build_tree(PLANE_V, PLANE_F, rng)
    bodies = np.stack([body_record(fut_trans=(0, 0, 1.0), static=True), body_record(fut_trans=(0, 0, 1.105))])
    sc = d2o.Scene([t_a, t_b], [0, 1], bodies)
    off, ints, reals = sc.narrow([[0, 1]], filter=False)
    assert off[1] == 8
    off, ints, reals = sc.narrow([[0, 1]], filter=True)
    assert off[1] == 1


# --- collision_detection/test_continuous_comparison.cpp --------------------
@pytest.mark.parametrize("prec", PRECS)
def test_sphere_sphere(prec):
    # test_continuous_comparison.cpp:6-32
    f = lambda *a: d2o.sphere_sphere(*a, prec=prec)
    assert f([0, 0, 0], [1, 0, 0], 1.0, [0, -0.5, 0], [-1, 0, 0], 1.0) == 0.0
    assert f([0, 0, 0], [1, 0, 0], 0.5, [0, -1.001, 0], [-1, 0, 0], 0.5) == -1.0
    assert f([0, 0, 0], [1, 0, 0], 0.5, [0, -1.001, 0], [1, 0.1, 0], 0.5) == pytest.approx(0.01, rel=1e-4)


A_T_START = np.array([[0.9961947202682495, -0.0858316496014595, -0.015134435147047043, 0.0],
                      [0.08715573698282242, 0.981060266494751, 0.17298738658428192, 0.0],
                      [0.0, -0.1736481785774231, 0.9848077297210693, 0.0], [0, 0, 0, 1.0]])


def _m(rows):
    return np.array(rows + [[0, 0, 0, 1.0]])


@pytest.mark.parametrize("prec", PRECS)
def test_ccd_no_intersect_far(prec):
    # test_continuous_comparison.cpp:34-75
    a_te = _m([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 2.0]])
    b_ts = _m([[0.912834644317627, -0.24342688918113708, -0.32783564925193787, -0.8726963400840759],
               [0.40832462906837463, 0.548050582408905, 0.7300079464912415, 1.8028504848480225],
               [0.0019669521134346724, -0.8002399206161499, 0.5996767282485962, 4.524521350860596]])
    b_te = _m([[0.9537079334259033, 0.2651883065700531, -0.14183202385902405, 0.0],
               [-0.05004240944981575, 0.6049837470054626, 0.7946637272834778, 1.0],
               [0.29654160141944885, -0.7507795095443726, 0.5902484655380249, 3.9175784587860107]])
    d, P, Q, toc = d2o.tri_tri_ccd(TRI1, A_T_START, a_te, TRI1, b_ts, b_te, prec)
    Pw, Qw = np.array([-0.951562, -0.004578, 2.0]), np.array([-0.954302, 1.05403, 3.62159])
    assert toc == 1.0
    assert np.all(np.abs(P - Pw) < 0.1) and np.all(np.abs(Q - Qw) < 0.1)
    assert abs(d - np.linalg.norm(Pw - Qw)) < 0.05


@pytest.mark.parametrize("prec", PRECS)
def test_ccd_no_intersect_close(prec):
    # test_continuous_comparison.cpp:77-117
    a_te = _m([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0.43169140815734863]])
    b_ts = _m([[0.912834644317627, -0.24342688918113708, -0.32783564925193787, 1.958876609802246],
               [0.40832462906837463, 0.548050582408905, 0.7300079464912415, -0.9190752506256104],
               [0.0019669521134346724, -0.8002399206161499, 0.5996767282485962, -1.093824863433838]])
    b_te = _m([[0.9537079334259033, 0.2651883065700531, -0.14183202385902405, 1.958876609802246],
               [-0.05004240944981575, 0.6049837470054626, 0.7946637272834778, -0.9190752506256104],
               [0.29654160141944885, -0.7507795095443726, 0.5902484655380249, 1.0489931106567383]])
    d, P, Q, toc = d2o.tri_tri_ccd(TRI1, A_T_START, a_te, TRI1, b_ts, b_te, prec)
    Pw, Qw = np.array([0.64, -0.73, 0.38]), np.array([1.0, -0.95, 0.44])
    assert abs(toc - 0.79) < 0.05
    assert np.all(np.abs(P - Pw) < 0.05) and np.all(np.abs(Q - Qw) < 0.05)
    assert abs(d - np.linalg.norm(Pw - Qw)) < 0.05


@pytest.mark.parametrize("prec", PRECS)
def test_ccd_intersect(prec):
    # test_continuous_comparison.cpp:119-160
    a_te = _m([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 2.0]])
    b_ts = _m([[0.9396926164627075, 0.1710100769996643, 0.29619812965393066, 0.0],
               [0.0, 0.8660253882408142, -0.5, 1.0],
               [-0.3420201241970062, 0.46984633803367615, 0.813797652721405, 1.2651395797729492]])
    b_te = _m([[0.9396926164627075, 0.0, 0.3420201241970062, 0.0], [0.0, 1.0, 0.0, 1.0],
               [-0.3420201241970062, 0.0, 0.9396926164627075, 1.5]])
    d, P, Q, toc = d2o.tri_tri_ccd(TRI1, A_T_START, a_te, TRI1, b_ts, b_te, prec)
    W = np.array([-0.265017, -0.843225, 0.591865])
    assert abs(toc - 0.24) < 0.05
    assert np.all(np.abs(P - W) < 0.1) and np.all(np.abs(Q - W) < 0.1)
    assert d == 0.0


TRI_EE = np.array([1, -1, 0, -1, 1, 0, -1, -1, 0], dtype=float)
A_EUL = [deg(66.2817), deg(-45.7037), deg(-111.169)]
B_EUL = [deg(261.326), deg(52.4147), deg(-101.34)]


def _ab_edge(M):
    t = transformed(TRI_EE, M, "f64")
    return t[0:6]


@pytest.mark.parametrize("prec", PRECS)
def test_edge_edge_newton_intersect(prec):
    # test_continuous_comparison.cpp:163-190
    a_s = _ab_edge(fixture_pose(A_EUL, [0, 0, 0]))
    a_e = _ab_edge(fixture_pose(A_EUL, [0, 0, 0.7]))
    b_s = _ab_edge(fixture_pose(B_EUL, [0, 0, 1.3]))
    b_e = _ab_edge(fixture_pose(B_EUL, [0, 0, 0.5]))
    d, P, Q, toc = d2o.edge_edge_ccd(a_s, a_e, b_s, b_e, prec)
    assert abs(toc - 0.87) < 0.01
    assert np.all(np.abs(P - [0, 0, 0.6]) < 0.01) and np.all(np.abs(Q - [0, 0, 0.6]) < 0.01)
    assert abs(d) < 0.01


@pytest.mark.parametrize("prec", PRECS)
def test_edge_edge_newton_no_intersect(prec):
    # test_continuous_comparison.cpp:192-219
    a_s = _ab_edge(fixture_pose(A_EUL, [0, 0, 0]))
    a_e = _ab_edge(fixture_pose(A_EUL, [0, 0, 0.5]))
    b_s = _ab_edge(fixture_pose(B_EUL, [0, 0, 1.3]))
    b_e = _ab_edge(fixture_pose(B_EUL, [0, 0, 0.7]))
    d, P, Q, toc = d2o.edge_edge_ccd(a_s, a_e, b_s, b_e, prec)
    assert abs(toc - 1.0) < 0.02
    assert np.all(np.abs(P - [0, 0, 0.5]) < 0.1) and np.all(np.abs(Q - [-0.01, -0.11, 0.68]) < 0.1)
    assert abs(d - 0.2) < 0.05


def test_continuous_bucket_pyramid_plane():
    # test_continuous_comparison.cpp:222-263: one filtered contact, toc ~ 0.9
    # (plane z=2-t meets the apex z=1.1 at t=0.9).  The plane's end pose
    # (0,0,1) is reached by velocity (0,0,-1) over max_time 1.
    rng = d2o.pcg_seed(42, 54)
    t_a = d2o.build_tree(PYRAMID_V, PYRAMID_F, rng)
    t_b = d2o.build_tree(PLANE_V, PLANE_F, rng)
    bodies = np.stack([body_record(trans=(0, 0, 1.0), static=True), body_record(trans=(0, 0, 2.0), vel=(0, 0, -1.0))])
    sc = d2o.Scene([t_a, t_b], [0, 1], bodies)
    off, ints, reals, per_pair = sc.ccd([[0, 1]], [1.0])
    after = sc.bodies()
    assert np.allclose(after[1, 18:21], [0, 0, 1.0])
    ints, reals = sc.filter_last(continuous=True)
    assert ints.shape[0] == 1
    assert reals[0, 10] == pytest.approx(0.9, rel=1.2e-5)
