"""Independent pins of the oracle's functions that no reference known-answer
test covers (CPU only).

Each test checks the oracle against something that does not share its code:
hand-computed branch values, closed forms, brute-force sampling, Newton's
method re-derived from the path it minimises, or a depth-first restatement of
the traversals (tests/traversal_restated.py).  A plausible slip -- a dropped
term, a swapped side, a wrong branch, a rotation multiplied on the wrong side
-- fails at least one of them.

  step rule      separate_clusters.cpp:496-564: one hand case per branch
  cluster step   separate_clusters.cpp:570-598 and the clamp of
                 time_step_selection_dynamic_continuous.cpp:19
  extrapolate    state.cpp:90-102, basic_utils.cpp:36-44: closed forms
  futurePointVelocity  particle.cpp:72-78: test_particle.cpp:22-40 + closed form
  edgeEdgeCCD    continuous_comparison.h:565-673: brute force in time
  traversals     full_tree_comparison.h:13-219: depth-first restatement and
                 the all-leaf-pairs superset
"""
import math

import numpy as np
import pytest

import d2o
import traversal_restated as tr
from delta2_amd import scenes

DT32 = float(np.float32(1.0 / 60.0))


# --- helpers -----------------------------------------------------------------
def quat_mul(p, q):
    """Hamilton product of (w, x, y, z) quaternions."""
    w1, x1, y1, z1 = p
    w2, x2, y2, z2 = q
    return np.array([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2, w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                     w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2, w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2])


def quat_rot(q):
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def axis_quat(axis, angle):
    a = np.asarray(axis, dtype=float)
    a = a / np.linalg.norm(a)
    return np.concatenate([[math.cos(angle / 2)], math.sin(angle / 2) * a])


def record(q=(1, 0, 0, 0), trans=(0, 0, 0), vel=(0, 0, 0), L=(0, 0, 0), inv_inertia=np.eye(3), last=1.0,
           static=False, geo_eps=0.01):
    """A body record whose current and future states are equal."""
    r = np.zeros(40)
    r[0:4] = q
    r[4:7] = trans
    r[7:10] = vel
    r[10:13] = L
    r[14:28] = r[0:14]
    r[28:37] = np.asarray(inv_inertia, dtype=float).ravel()
    r[37] = geo_eps
    r[38] = 1.0 if static else 0.0
    r[39] = last
    return r


def contact(A, B, toc, eps_a=0.01, eps_b=0.01):
    return np.concatenate([A, B, [eps_a, eps_b, 0.0, 0.0, toc]])


# --- futurePointVelocity ------------------------------------------------------
def test_future_point_velocity_reference_vector():
    """test_particle.cpp:22-40: the unit-density cube [-1,1]^3 (mass 8,
    I = 16/3 Id), L = (1,0,0): the velocity at (0,0,-1) points along +y, with
    and without a linear velocity (0,1,0)."""
    cube_inv = np.eye(3) * 3.0 / 16.0
    b = record(L=(1, 0, 0), inv_inertia=cube_inv)
    v = d2o.future_point_velocity(b, [0, 0, -1])
    assert abs(v @ [0, 1, 0] / np.linalg.norm(v) - 1.0) < 1e-12
    assert np.allclose(v, [0, 3.0 / 16.0, 0], rtol=0, atol=1e-16)
    b = record(vel=(0, 1, 0), L=(1, 0, 0), inv_inertia=cube_inv)
    v = d2o.future_point_velocity(b, [0, 0, -1])
    assert abs(v @ [0, 1, 0] / np.linalg.norm(v) - 1.0) < 1e-12
    assert np.allclose(v, [0, 1.0 + 3.0 / 16.0, 0], rtol=0, atol=1e-15)


def test_future_point_velocity_closed_form():
    """v + (R I_b^-1 R^T L) x (p - c) of the future state, for rotated
    anisotropic bodies; the current state must not enter."""
    rng = np.random.default_rng(3)
    for _ in range(50):
        q = scenes.random_rotation(rng)
        Ib = np.diag(rng.uniform(0.2, 3.0, 3))
        r = record(q=q, trans=rng.normal(size=3), vel=rng.normal(size=3), L=rng.normal(size=3), inv_inertia=Ib)
        r[0:14] = 0.0  # garbage current state
        p = rng.normal(size=3)
        R = quat_rot(q)
        w = R @ Ib @ R.T @ r[24:27]
        want = r[21:24] + np.cross(w, p - r[18:21])
        assert np.allclose(d2o.future_point_velocity(r, p), want, rtol=0, atol=1e-13)


# --- extrapolate -----------------------------------------------------------------
def test_extrapolate_without_angular_momentum():
    s = record(q=axis_quat([1, 2, 3], 0.7), trans=(1, 2, 3), vel=(0.5, -1, 2))[0:14]
    s[13] = 4.0
    o = d2o.extrapolate(s, np.eye(3), 0.25)
    assert np.array_equal(o[0:4], s[0:4])  # exp(0) is the identity: rotation unchanged, bit for bit
    assert np.array_equal(o[4:7], s[4:7] + 0.25 * s[7:10])
    assert np.array_equal(o[7:13], s[7:13]) and o[13] == 4.25


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_extrapolate_principal_axis(axis):
    """Identity pose, diagonal I^-1, L along a principal axis: a rotation by
    I^-1_kk L t about that axis."""
    inv = np.diag([0.5, 2.0, 3.0])
    L = np.zeros(3)
    L[axis] = 1.7
    s = record(L=L, inv_inertia=inv)[0:14]
    t = 0.3
    o = d2o.extrapolate(s, inv, t)
    e = np.zeros(3)
    e[axis] = 1.0
    assert np.allclose(o[0:4], axis_quat(e, inv[axis, axis] * 1.7 * t), rtol=0, atol=1e-15)


def test_extrapolate_multiplies_the_delta_on_the_right():
    """state.cpp:95-99: omega = R I_b^-1 R^T L (world frame), then
    rotation * exp(omega t / 2) -- the delta on the right, although omega is
    a world-frame vector.  Closed form for rotated anisotropic bodies; the
    left product differs, so a swapped order fails."""
    rng = np.random.default_rng(5)
    for _ in range(40):
        q = scenes.random_rotation(rng)
        inv = np.diag(rng.uniform(0.2, 3.0, 3))
        L = rng.normal(size=3)
        t = rng.uniform(0.001, 0.5)
        s = record(q=q, L=L, inv_inertia=inv)[0:14]
        R = quat_rot(q)
        w = R @ inv @ R.T @ L
        h = 0.5 * w * t
        n = np.linalg.norm(h)
        delta = np.concatenate([[math.cos(n)], math.sin(n) * h / n])
        o = d2o.extrapolate(s, inv, t)
        assert np.allclose(o[0:4], quat_mul(q, delta), rtol=0, atol=1e-14)
        assert not np.allclose(quat_mul(delta, q), quat_mul(q, delta), atol=1e-6)


# --- the TOI -> step rule -----------------------------------------------------------
# One contact per case between body a (moving) and body b (at rest), n = A - B,
# eps_a + eps_b = 0.02, max_time 0.1; expected scales worked out by hand from
# separate_clusters.cpp:509-550 with lerp(x, y, s) = x (1 - s) + y s.
STEP_CASES = [
    # touching (|n| < 1e-4) with toc == 0: skipped (:511-514)
    ("touching toc=0", [0, 0, 0], [0, 0, 0], 0.0, [0.01, 0, 0], 1.0),
    # touching, toc 0.5, |v| 0.01: start = 0.5 - 0.01*0.1/0.02 = 0.45, lerp(0.45, 0.5, 0.75) (:515-519)
    ("touching", [0, 0, 0], [0, 0, 0], 0.5, [0.01, 0, 0], 0.4875),
    # touching, |v| 1: start = max(0, 0.5 - 5) = 0, lerp(0, 0.5, 0.75)
    ("touching start clamped", [0, 0, 0], [0, 0, 0], 0.5, [1, 0, 0], 0.375),
    # approaching, first entry (:531-537): proj_vel 2, depth 0.02-0.015, proj_dist 2*0.4*0.1;
    # lerp(0.4 (1 - 0.005/0.08), 0.4, 0.5)
    ("first entry", [0.015, 0, 0], [0, 0, 0], 0.4, [-2, 0, 0], 0.3875),
    # already inside, future point beyond 1 (:542-548): proj_dist 0.0004 <= depth 0.005,
    # future 0.4 + 0.6*0.02/0.001 = 12.4 -> lerp(0.4, 1, 0.8)
    ("already inside, capped", [0.015, 0, 0], [0, 0, 0], 0.4, [-0.01, 0, 0], 0.88),
    # already inside: |n| 0.001, depth 0.019, proj_vel 0.4, proj_dist 0.016;
    # future 0.4 + 0.6*0.02/0.04 = 0.7 -> lerp(0.4, 0.7, 0.8)
    ("already inside", [0.001, 0, 0], [0, 0, 0], 0.4, [-0.4, 0, 0], 0.64),
    # proj_dist = 1e-6*0.4*0.1 <= 1e-6: no change (:531)
    ("proj_dist below 1e-6", [0.015, 0, 0], [0, 0, 0], 0.4, [-1e-6, 0, 0], 1.0),
    # approaching but toc == 1: no change (:525)
    ("toc 1", [0.015, 0, 0], [0, 0, 0], 1.0, [-2, 0, 0], 1.0),
    # separating (n.v > 0, |n| >= 1e-4): no change (:509)
    ("separating", [0.015, 0, 0], [0, 0, 0], 0.4, [2, 0, 0], 1.0),
    # no relative motion: normalized(0) = 0, so n.v = 0 <= 0 enters the approaching
    # branch with proj_vel 0 and changes nothing
    ("at rest", [0.015, 0, 0], [0, 0, 0], 0.4, [0, 0, 0], 1.0),
]


@pytest.mark.parametrize("name,A,B,toc,vel,want", STEP_CASES, ids=[c[0] for c in STEP_CASES])
def test_step_rule_branch(name, A, B, toc, vel, want):
    a = record(vel=vel)
    b = record()
    s = d2o.step_scale(contact(A, B, toc), a, b, 0.1)
    assert abs(s - want) <= 1e-12, (name, s, want)


def test_step_rule_takes_the_smallest_scale():
    cs = np.concatenate([contact([0.001, 0, 0], [0, 0, 0], 0.4),  # 0.64 with v = -0.4
                         contact([0.015, 0, 0], [0, 0, 0], 0.4),  # 0.3875 at v = -2 but here v = -0.4
                         contact([0, 0, 0], [0, 0, 0], 0.5)])
    a = record(vel=[-0.4, 0, 0])
    # second contact at v = -0.4: proj_dist 0.016 > depth 0.005 -> lerp(0.4 (1 - 0.005/0.016), 0.4, 0.5) = 0.3375
    # third (touching, |v| 0.4): start = max(0, 0.5 - 2) = 0 -> 0.375
    assert abs(d2o.step_scale(cs, a, record(), 0.1) - 0.3375) <= 1e-12


def test_step_rule_relative_velocity_from_rotation():
    """The relative velocity comes from futurePointVelocity: body a spinning
    about z at 2 rad/s moves its point (0,1,0) at (-2,0,0); with b at rest at
    n = (0.015,0,0) below it this is the 'first entry' case (0.3875)."""
    a = record(L=(0, 0, 2.0))  # I^-1 = Id: omega = L
    b = record()
    A = np.array([0.0, 1.0, 0.0])
    s = d2o.step_scale(contact(A, A - [0.015, 0, 0], 0.4), a, b, 0.1)
    assert abs(s - 0.3875) <= 1e-12
    # the same motion given to b with the opposite sign
    s = d2o.step_scale(contact(A, A - [0.015, 0, 0], 0.4), record(), record(vel=(2, 0, 0)), 0.1)
    assert abs(s - 0.3875) <= 1e-12


# --- cluster step: minimum, first-pair read, clamp ----------------------------------------
def _two_leaf_scene(records, tree_of_body=None):
    """Bodies over the pyramid (tree 0) and plane (tree 1) meshes of
    test_comparison.cpp:93-105 and primitive_geo.cpp."""
    from test_oracle_kat import PLANE_F, PLANE_V, PYRAMID_F, PYRAMID_V
    rng = d2o.pcg_seed(42, 54)
    trees = [d2o.build_tree(PYRAMID_V, PYRAMID_F, rng), d2o.build_tree(PLANE_V, PLANE_F, rng)]
    tob = [i % 2 for i in range(len(records))] if tree_of_body is None else tree_of_body
    return d2o.Scene(trees, tob, np.stack(records))


def test_cluster_step_clamp_and_first_pair():
    far = lambda i, **k: record(trans=(10.0 * i, 0, 0), **k)  # noqa: E731 -- no CCD contact
    # no contacts: the steps stay the bodies' last steps; the first pair reads body a
    sc = _two_leaf_scene([far(0, last=0.005), far(1, last=0.01)])
    assert sc.cluster_step([[0, 1]], np.float32(1 / 60)) == 0.005
    # clamp below (1e-3) and above (the float dt)
    sc = _two_leaf_scene([far(0, last=1e-4), far(1, last=1e-4)])
    assert sc.cluster_step([[0, 1]], np.float32(1 / 60)) == 1e-3
    sc = _two_leaf_scene([far(0, last=1.0), far(1, last=1.0)])
    assert sc.cluster_step([[0, 1]], np.float32(1 / 60)) == DT32
    # first pair with a static body a: the cluster starts from step_size[a_id]
    # (separate_clusters.cpp:589-590), the static body's last step, although
    # only the dynamic body steps
    sc = _two_leaf_scene([far(0, last=0.5, static=True), far(1, last=0.004)])
    assert sc.cluster_step([[0, 1]], np.float32(1 / 60)) == DT32
    # later pairs of two dynamic bodies contribute body a's step (:576-577, :593)
    sc = _two_leaf_scene([far(0, last=0.5, static=True), far(1, last=0.004), far(2, last=0.003)])
    assert sc.cluster_step([[0, 1], [1, 2]], np.float32(1 / 60)) == 0.004


def test_cluster_step_with_contacts_is_scale_times_interval():
    """A plane (body a, descending at 1 m/s) over a static pyramid whose apex
    it passes through at t = 0.5 of a 0.01 s interval
    (test_continuous_comparison.cpp:222-263 geometry): touching contacts at
    toc 0.5, |v| = 1, so scale = lerp(max(0, 0.5 - 0.01/0.02), 0.5, 0.75)
    = 0.375, and the cluster step is 0.01 x 0.375."""
    plane = record(trans=(0, 0, 1.105), vel=(0, 0, -1.0), last=0.01)
    pyramid = record(trans=(0, 0, 1.0), static=True, last=0.01)
    sc = _two_leaf_scene([plane, pyramid], tree_of_body=[1, 0])
    off, ints, reals, pp = sc.ccd([[0, 1]], [0.01])
    assert off[1] > 0
    fut = [plane.copy(), pyramid.copy()]
    for r in fut:
        r[14:28] = d2o.extrapolate(r[0:14], r[28:37], 0.01)
    scale = d2o.step_scale(reals, fut[0], fut[1], 0.01)
    assert scale == pp[0, 0] and abs(scale - 0.375) <= 1e-12
    sc = _two_leaf_scene([plane, pyramid], tree_of_body=[1, 0])
    assert sc.cluster_step([[0, 1]], np.float32(1 / 60)) == 0.01 * scale


# --- edgeEdgeCCD by brute force in time -------------------------------------------------
def segdist(p0, p1, q0, q1):
    """Closest points of two segments (Ericson, Real-Time Collision Detection 5.1.9)."""
    d1, d2, r = p1 - p0, q1 - q0, p0 - q0
    a, e, f = d1 @ d1, d2 @ d2, d2 @ r
    c, b = d1 @ r, d1 @ d2
    den = a * e - b * b
    s = min(max((b * f - c * e) / den, 0.0), 1.0) if den > 1e-14 * a * e else 0.0
    t = (b * s + f) / e
    if t < 0.0:
        t, s = 0.0, min(max(-c / a, 0.0), 1.0)
    elif t > 1.0:
        t, s = 1.0, min(max((b - c) / a, 0.0), 1.0)
    P, Q = p0 + s * d1, q0 + t * d2
    return np.linalg.norm(P - Q), P, Q


def segdist_many(p0, p1, q0, q1):
    """segdist over arrays of segments (rows), vectorised for the time sweeps."""
    d1, d2, r = p1 - p0, q1 - q0, p0 - q0
    a, e, f = (d1 * d1).sum(1), (d2 * d2).sum(1), (d2 * r).sum(1)
    c, b = (d1 * r).sum(1), (d1 * d2).sum(1)
    den = a * e - b * b
    with np.errstate(divide="ignore", invalid="ignore"):
        s = np.where(den > 1e-14 * a * e, np.clip((b * f - c * e) / den, 0.0, 1.0), 0.0)
        t = (b * s + f) / e
        s = np.where(t < 0.0, np.clip(-c / a, 0.0, 1.0), np.where(t > 1.0, np.clip((b - c) / a, 0.0, 1.0), s))
        t = np.clip(t, 0.0, 1.0)
    return np.linalg.norm(p0 + s[:, None] * d1 - (q0 + t[:, None] * d2), axis=1)


def at(e0, e1, t):
    """Endpoints of an edge lerped between its start and end positions."""
    return e0[:3] * (1 - t) + e1[:3] * t, e0[3:] * (1 - t) + e1[3:] * t


def point_seg_dist(p, a, b):
    ab = b - a
    s = min(max((p - a) @ ab / (ab @ ab), 0.0), 1.0)
    return np.linalg.norm(p - (a + s * ab))


def random_motion(rng, scale=1.0):
    a0 = rng.normal(size=6)
    b0 = rng.normal(size=6)
    return a0, a0 + scale * rng.normal(size=6), b0, b0 + scale * rng.normal(size=6)


def coplanarity(as_, ae, bs, be, t):
    a0, a1 = at(as_, ae, t)
    b0, b1 = at(bs, be, t)
    return (b1 - a0) @ np.cross(a1 - a0, b0 - a0)


def test_edge_edge_result_is_a_real_configuration():
    """P lies on edge a and Q on edge b at the returned time, |P - Q| is the
    returned distance and equals the segments' true distance then (Ericson's
    closest points), and nothing over the interval is closer than the true
    minimum (sampled, with the motion's Lipschitz slack)."""
    rng = np.random.default_rng(11)
    for k in range(300):
        as_, ae, bs, be = random_motion(rng)
        d, P, Q, toc = d2o.edge_edge_ccd(as_, ae, bs, be)
        assert 0.0 <= toc <= 1.0
        a0, a1 = at(as_, ae, toc)
        b0, b1 = at(bs, be, toc)
        assert point_seg_dist(P, a0, a1) <= 1e-12 and point_seg_dist(Q, b0, b1) <= 1e-12, k
        assert abs(np.linalg.norm(P - Q) - d) <= 1e-15 * max(d, 1.0)
        assert abs(segdist(a0, a1, b0, b1)[0] - d) <= 1e-12, k
        ts = np.linspace(0.0, 1.0, 2001)[:, None]
        lo = segdist_many(*at(as_, ae, ts), *at(bs, be, ts)).min()
        speed = max(np.linalg.norm(ae[:3] - as_[:3]), np.linalg.norm(ae[3:] - as_[3:])) + \
            max(np.linalg.norm(be[:3] - bs[:3]), np.linalg.norm(be[3:] - bs[3:]))
        assert d >= lo - speed * 0.5 / 2000 - 1e-12, k
        if coplanarity(as_, ae, bs, be, 0.0) * coplanarity(as_, ae, bs, be, 1.0) >= 0.0:
            # no sign change: Newton's first iterate is t = 0 and the smallest iterate is kept
            assert d <= segdist(*at(as_, ae, 0.0), *at(bs, be, 0.0))[0] + 1e-15, k


def test_edge_edge_crossing_time_by_bisection():
    """Segments that pass through each other at a known time t*: the
    coplanarity sign change is bracketed by 20 bisection steps
    (continuous_comparison.h:574-598), so toc is within 2^-20 of t* (or
    exactly where |dot| < 1e-8) and the distance there is tiny."""
    rng = np.random.default_rng(12)
    for _ in range(200):
        t_star = rng.uniform(0.05, 0.95)
        a0, a1 = rng.normal(size=3), rng.normal(size=3)
        ua, ub = rng.uniform(0.2, 0.8, size=2)
        x = a0 + ua * (a1 - a0)  # the crossing point on edge a
        db = rng.normal(size=3)
        b0x, b1x = x - ub * db, x + (1 - ub) * db  # edge b through x at t*
        v = rng.normal(size=3)  # b translates with velocity v
        bs = np.concatenate([b0x - t_star * v, b1x - t_star * v])
        be = bs + np.tile(v, 2)
        as_ = ae = np.concatenate([a0, a1])
        d, P, Q, toc = d2o.edge_edge_ccd(as_, ae, bs, be)
        assert abs(toc - t_star) <= 2.0 ** -20 + 1e-12
        assert d <= (np.linalg.norm(v) + 1.0) * 2.0 ** -19


def _newton_from_its_meaning(as_, ae, bs, be):
    """The Newton phase of edgeEdgeCCD (continuous_comparison.h:612-637)
    derived from what it minimises rather than from its sympy expression: at
    each iterate t the closest points fix pa, pb (their parameters on the
    edges at t), and the step minimises f(s) = |D(s)| with
      D(s) = pa (a_end.B + a_start.B (1-s)) - pb (b_end.B + b_start.B (1-s))
             + (1-pa) lerp(a_start.A, a_end.A, s) - (1-pb) lerp(b_start.A, b_end.A, s)
    (the B endpoints as the expression writes them, :630).  D is affine in s,
    so f' = D.D'/|D| and f'' = D'.D'/|D| - (D.D')^2/|D|^3 in closed form;
    t <- clamp01(t - f'/f'')."""
    aA0, aB0, aA1, aB1 = as_[:3], as_[3:], ae[:3], ae[3:]
    bA0, bB0, bA1, bB1 = bs[:3], bs[3:], be[:3], be[3:]
    best, best_t = np.inf, 0.0
    t = 0.0
    for _ in range(10):
        a0, a1 = at(as_, ae, t)
        b0, b1 = at(bs, be, t)
        dist, P, Q = segdist(a0, a1, b0, b1)
        if dist < best:
            best, best_t = dist, t
        pa = (P - a0) @ (a1 - a0) / ((a1 - a0) @ (a1 - a0))
        pb = (Q - b0) @ (b1 - b0) / ((b1 - b0) @ (b1 - b0))

        def D(s):
            return (pa * (aB1 + aB0 * (1 - s)) - pb * (bB1 + bB0 * (1 - s)) + (1 - pa) * (aA0 * (1 - s) + aA1 * s)
                    - (1 - pb) * (bA0 * (1 - s) + bA1 * s))

        Dt, dD = D(t), D(1.0) - D(0.0)
        n = np.linalg.norm(Dt)
        f1 = Dt @ dD / n
        f2 = dD @ dD / n - (Dt @ dD) ** 2 / n ** 3
        t = min(max(t - f1 / f2, 0.0), 1.0)
    return best, best_t


def test_edge_edge_newton_matches_its_derivation():
    """Cases without a coplanarity sign change run only the Newton phase: its
    best iterate must be the one the re-derived Newton iteration finds."""
    rng = np.random.default_rng(13)
    checked = 0
    while checked < 300:
        as_, ae, bs, be = random_motion(rng, scale=0.5)
        if coplanarity(as_, ae, bs, be, 0.0) * coplanarity(as_, ae, bs, be, 1.0) < 0.0:
            continue
        d, P, Q, toc = d2o.edge_edge_ccd(as_, ae, bs, be)
        want_d, want_t = _newton_from_its_meaning(as_, ae, bs, be)
        assert abs(d - want_d) <= 1e-9 * max(1.0, want_d), (checked, d, want_d)
        assert abs(toc - want_t) <= 1e-7, (checked, toc, want_t)
        checked += 1


# --- traversals: depth-first restatement and all-leaf-pairs superset -------------------------
def _scene_pairs():
    """Pairs of multi-level trees with CCD contacts: subdivision-2 icospheres
    (dense spinning scene), ellipsoid hulls, and the 4096-body workload's
    generator (subdivision-3 icospheres and hulls) -- per scene its busiest
    pair and pairs at the 75th and 50th percentile of the contact count."""
    from gpu_cases import dense_scene, hull_scene, sparse_scene
    out = []
    for make, steps in ((dense_scene, 2), (hull_scene, 2), (sparse_scene, 16)):
        s = make()
        rec = s.records
        osc = d2o.Scene(s.trees, s.tree_of_body, rec)
        mt = steps * scenes.DT
        off, _, _, _ = osc.ccd(s.pairs, np.full(s.pairs.shape[0], mt))
        n = np.diff(off)
        hit = np.nonzero(n > 0)[0]
        order = hit[np.argsort(-n[hit], kind="stable")]
        pick = sorted({int(order[0]), int(order[len(order) // 4]), int(order[len(order) // 2])})
        out += [(s, p, mt) for p in pick]
    return out


@pytest.fixture(scope="module")
def traversal_cases():
    return _scene_pairs()


def _bodies(s, p, records):
    a, b = s.pairs[p]
    ta = s.trees[s.tree_of_body[a]]
    tb = s.trees[s.tree_of_body[b]]
    return a, b, tr.Body(ta, records[a]), tr.Body(tb, records[b])


def test_discrete_traversal_matches_depth_first_restatement(traversal_cases):
    from gpu_cases import with_future
    seen = 0
    for s, p, mt in traversal_cases:
        rec = with_future(s.records, mt)
        a, b, A, B = _bodies(s, p, rec)
        osc = d2o.Scene(s.trees, s.tree_of_body, rec)
        off, ints, reals = osc.narrow([[a, b]], filter=False)
        want_i, want_r = tr.compare_trees_full(A, B)
        assert np.array_equal(ints[:, 1:], want_i[:, 1:]), (s.name, p)
        assert np.array_equal(reals, want_r), (s.name, p)
        seen += ints.shape[0]
        # inner-level node pairs expanded: a multi-level traversal was exercised
        assert ints.shape[0] == 0 or ints[:, 6].max() >= 2
    assert seen > 20


def test_ccd_traversal_matches_depth_first_restatement(traversal_cases):
    seen = 0
    for s, p, mt in traversal_cases:
        a, b, A, B = _bodies(s, p, s.records)
        osc = d2o.Scene(s.trees, s.tree_of_body, s.records)
        off, ints, reals, pp = osc.ccd([[a, b]], [mt])
        want_i, want_r, want_min = tr.compare_trees_full_continuous(A, B, mt)
        assert np.array_equal(ints[:, 1:], want_i[:, 1:]), (s.name, p)
        assert np.array_equal(reals, want_r), (s.name, p)
        assert pp[0, 1] == want_min
        fut = [s.records[a].copy(), s.records[b].copy()]
        for r in fut:
            r[14:28] = d2o.extrapolate(r[0:14], r[28:37], mt)
        want_scale = d2o.step_scale(want_r, fut[0], fut[1], mt) if want_r.shape[0] else 1.0
        assert pp[0, 0] == want_scale
        seen += ints.shape[0]
    assert seen > 20


def test_traversal_contacts_are_leaf_pair_contacts(traversal_cases):
    """Every contact a traversal emits is a contact of its leaf pair compared
    directly (same points, eps and toc), and the traversal finds most of what
    comparing all leaf pairs finds (the surrogates' eps bound is over the
    vertices only, construct_surrogate_tree.h:172-185, so it may miss some)."""
    from gpu_cases import with_future
    found = total = 0
    for s, p, mt in traversal_cases:
        # all leaf pairs of two subdivision-3 trees through the connected CCD
        # comparison take minutes; those pairs are checked in the discrete phase
        for continuous in ((False,) if s.name == "c2" else (False, True)):
            rec = s.records if continuous else with_future(s.records, mt)
            a, b = s.pairs[p]
            osc = d2o.Scene(s.trees, s.tree_of_body, rec)
            if continuous:
                _, ints, reals, _ = osc.ccd([[a, b]], [mt])
            else:
                _, ints, reals = osc.narrow([[a, b]], filter=False)
            ex_i, ex_r = d2o.Scene(s.trees, s.tree_of_body, rec).exhaustive_leaves(a, b, continuous, mt)
            table = {tuple(r[1:6]): k for k, r in enumerate(ex_i)}
            for k in range(ints.shape[0]):
                key = tuple(ints[k, 1:6])
                assert key in table, (s.name, p, continuous, key)
                assert np.array_equal(reals[k], ex_r[table[key]])
            found += ints.shape[0]
            total += ex_i.shape[0]
    assert found >= 0.5 * total, (found, total)
