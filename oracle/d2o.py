"""ctypes binding of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

The oracle is a C++ restatement of Delta2's contact-detection path
(collision_detection/*.h, common/*.h, model/* of the reference).  Only the
test-suite, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` load it, and only as the checker.  The product never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "build", "libd2oracle.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_vp = ctypes.c_void_p

_lib = None


def build():
    """Compile the oracle library (g++, -O2 -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        for prec in ("f64", "f32"):
            for name, args in [
                ("transformation_matrix_euler", [_dp, _dp, _dp]),
                ("transform_point", [_dp, _dp, _dp]),
                ("comparison_dist", [_dp, _dp, _dp]),
                ("edge_closest", [_dp, _dp, _dp]),
                ("barycentric", [_dp, _dp, _dp]),
                ("intersect_segment", [_dp, _dp, _dp]),
                ("project_to_plane", [_dp, _dp, _dp]),
                ("point_tri_ccd", [_dp] * 7),
                ("edge_edge_ccd", [_dp] * 5),
                ("tri_tri_ccd", [_dp] * 7),
            ]:
                f = getattr(L, f"d2o_{name}_{prec}")
                f.argtypes = args
                f.restype = None
            f = getattr(L, f"d2o_dist_to_point_{prec}")
            f.argtypes = [_dp, _dp]
            f.restype = ctypes.c_double
            f = getattr(L, f"d2o_sphere_sphere_{prec}")
            f.argtypes = [_dp, _dp, ctypes.c_double, _dp, _dp, ctypes.c_double]
            f.restype = ctypes.c_double
        L.d2o_extrapolate.argtypes = [_dp, _dp, ctypes.c_double, _dp]
        L.d2o_state_transformation.argtypes = [_dp, _dp]
        L.d2o_future_point_velocity.argtypes = [_dp, _dp, _dp]
        L.d2o_pcg_draws.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int, _u32p]
        L.d2o_pcg_seed.argtypes = [ctypes.c_uint64, ctypes.c_uint64, _u64p]
        L.d2o_tree_build.argtypes = [_dp, ctypes.c_int, _ip, ctypes.c_int, _u64p]
        L.d2o_tree_build.restype = _vp
        L.d2o_tree_num_nodes.argtypes = [_vp]
        L.d2o_tree_num_nodes.restype = ctypes.c_int
        L.d2o_tree_export.argtypes = [_vp, _ip, _dp]
        L.d2o_tree_import.argtypes = [_ip, _dp, ctypes.c_int]
        L.d2o_tree_import.restype = _vp
        L.d2o_tree_free.argtypes = [_vp]
        L.d2o_scene_create.argtypes = [ctypes.c_int, ctypes.POINTER(_vp), ctypes.c_int, _ip, _dp]
        L.d2o_scene_create.restype = _vp
        L.d2o_scene_free.argtypes = [_vp]
        L.d2o_scene_get_bodies.argtypes = [_vp, _dp]
        L.d2o_narrow.argtypes = [_vp, ctypes.c_int, _ip, ctypes.c_int]
        L.d2o_narrow.restype = ctypes.c_int
        L.d2o_ccd.argtypes = [_vp, ctypes.c_int, _ip, _dp]
        L.d2o_ccd.restype = ctypes.c_int
        L.d2o_scene_result.argtypes = [_vp, _ip, _ip, _dp, _dp]
        L.d2o_filter_order.argtypes = [_dp, ctypes.c_int, ctypes.c_double, _ip, _ip]
        L.d2o_filter_order.restype = ctypes.c_int
        L.d2o_exhaustive_leaves.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double]
        L.d2o_exhaustive_leaves.restype = ctypes.c_int
        L.d2o_cluster_step.argtypes = [_vp, ctypes.c_int, _ip, ctypes.c_float]
        L.d2o_cluster_step.restype = ctypes.c_double
        L.d2o_time_pass.argtypes = [_vp, ctypes.c_int, _ip, _dp, ctypes.c_int, ctypes.c_int, _dp]
        L.d2o_time_pass.restype = ctypes.c_double
        L.d2o_scene_filter_result.argtypes = [_vp, ctypes.c_int]
        L.d2o_scene_filter_result.restype = ctypes.c_int
        L.d2o_step_scale.argtypes = [_dp, ctypes.c_int, _dp, _dp, ctypes.c_double]
        L.d2o_step_scale.restype = ctypes.c_double
        _fp = ctypes.POINTER(ctypes.c_float)
        L.d2o_mesh_bounds.argtypes = [_dp, ctypes.c_int, _fp, _dp]
        L.d2o_broad_phase.argtypes = [ctypes.c_int, _dp, _dp, _fp, _dp, _ip, ctypes.c_int, ctypes.c_double,
                                      ctypes.c_int, _ip]
        L.d2o_broad_phase.restype = ctypes.c_int
        L.d2o_clusters.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int, _ip, _ip, _ip, _ip, _dp, _dp, _dp]
        L.d2o_clusters.restype = ctypes.c_int
        L.d2o_interpolate.argtypes = [_dp, _dp, ctypes.c_double, _dp]
        L.d2o_interpolate.restype = ctypes.c_int
        L.d2o_solver_props.argtypes = [_vp, _dp, _dp]
        L.d2o_solve_cluster.argtypes = [_vp, ctypes.c_int, _ip, ctypes.c_int, _ip, ctypes.c_double, _dp,
                                        ctypes.POINTER(ctypes.c_int)]
        L.d2o_solve_cluster.restype = ctypes.c_int
        L.d2o_solve_results.argtypes = [_vp, _dp]
        L.d2o_warm_get.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _dp]
        L.d2o_warm_get.restype = ctypes.c_int
        L.d2o_commit.argtypes = [_vp, _dp]
        L.d2o_scene_set_bodies.argtypes = [_vp, _dp]
        _lib = L
    return _lib


def _d(a, n=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel())
    if n is not None:
        assert a.size == n, (a.size, n)
    return a


def _p(a):
    return a.ctypes.data_as(_dp)


def _call(name, prec, *arrays, out=0):
    o = np.zeros(out)
    arrs = [_d(a) for a in arrays]
    getattr(lib(), f"d2o_{name}_{prec}")(*[_p(a) for a in arrs], _p(o))
    return o


# --- primitives (prec is "f64" or "f32") ------------------------------------
def transformation_matrix_euler(euler, pos, prec="f64"):
    return _call("transformation_matrix_euler", prec, euler, pos, out=16).reshape(4, 4)


def transform_point(p, M, prec="f64"):
    return _call("transform_point", prec, p, M, out=3)


def comparison_dist(tri_a, tri_b, prec="f64"):
    o = _call("comparison_dist", prec, tri_a, tri_b, out=7)
    return o[0], o[1:4], o[4:7]


def edge_closest(e, f, prec="f64"):
    o = _call("edge_closest", prec, e, f, out=7)
    return o[0], o[1:4], o[4:7]


def barycentric(tri, p, prec="f64"):
    return _call("barycentric", prec, tri, p, out=3)


def intersect_segment(tri, e, prec="f64"):
    o = _call("intersect_segment", prec, tri, e, out=6)
    return o[0], bool(o[1]), bool(o[2]), o[3:6]


def project_to_plane(tri, p, prec="f64"):
    o = _call("project_to_plane", prec, tri, p, out=4)
    return bool(o[0]), o[1:4]


def dist_to_point(tri, p, prec="f64"):
    return getattr(lib(), f"d2o_dist_to_point_{prec}")(_p(_d(tri)), _p(_d(p)))


def sphere_sphere(ap, av, ar, bp, bv, br, prec="f64"):
    return getattr(lib(), f"d2o_sphere_sphere_{prec}")(_p(_d(ap)), _p(_d(av)), ar, _p(_d(bp)), _p(_d(bv)), br)


def point_tri_ccd(pt, pTs, pTe, tri, tTs, tTe, prec="f64"):
    o = _call("point_tri_ccd", prec, pt, pTs, pTe, tri, tTs, tTe, out=8)
    return o[0], o[1:4], o[4:7], o[7]


def edge_edge_ccd(a_s, a_e, b_s, b_e, prec="f64"):
    o = _call("edge_edge_ccd", prec, a_s, a_e, b_s, b_e, out=8)
    return o[0], o[1:4], o[4:7], o[7]


def tri_tri_ccd(a, aTs, aTe, b, bTs, bTe, prec="f64"):
    o = _call("tri_tri_ccd", prec, a, aTs, aTe, b, bTs, bTe, out=8)
    return o[0], o[1:4], o[4:7], o[7]


def extrapolate(state14, inv_inertia9, t):
    o = np.zeros(14)
    lib().d2o_extrapolate(_p(_d(state14, 14)), _p(_d(inv_inertia9, 9)), float(t), _p(o))
    return o


def state_transformation(state14):
    o = np.zeros(16)
    lib().d2o_state_transformation(_p(_d(state14, 14)), _p(o))
    return o.reshape(4, 4)


def future_point_velocity(body40, pt):
    """Particle::futurePointVelocity of a body record at world point pt."""
    o = np.zeros(3)
    lib().d2o_future_point_velocity(_p(_d(body40, 40)), _p(_d(pt, 3)), _p(o))
    return o


def pcg_draws(s0, s1, bound, n):
    o = np.zeros(n, dtype=np.uint32)
    lib().d2o_pcg_draws(s0, s1, bound, n, o.ctypes.data_as(_u32p))
    return o


def pcg_seed(s0=42, s1=54):
    st = np.zeros(2, dtype=np.uint64)
    lib().d2o_pcg_seed(s0, s1, st.ctypes.data_as(_u64p))
    return st


# --- trees ------------------------------------------------------------------
def build_tree(V, F, rng_state):
    """Oracle surrogate tree of mesh (V, F); rng_state (uint64[2]) advances in place.

    Returns (nodes int32[n,8], buckets float64[n,88])."""
    V = _d(V)
    F = np.ascontiguousarray(np.asarray(F, dtype=np.int32).ravel())
    assert rng_state.dtype == np.uint64
    h = lib().d2o_tree_build(_p(V), V.size // 3, F.ctypes.data_as(_ip), F.size // 3, rng_state.ctypes.data_as(_u64p))
    n = lib().d2o_tree_num_nodes(h)
    nodes = np.zeros((n, 8), dtype=np.int32)
    buckets = np.zeros((n, 88))
    lib().d2o_tree_export(h, nodes.ctypes.data_as(_ip), _p(buckets))
    lib().d2o_tree_free(h)
    return nodes, buckets


class Scene:
    """Oracle scene: trees (flat arrays) + body records (40 doubles each)."""

    def __init__(self, trees, tree_of_body, bodies):
        L = lib()
        handles = []
        for nodes, buckets in trees:
            nodes = np.ascontiguousarray(nodes, dtype=np.int32)
            buckets = np.ascontiguousarray(buckets, dtype=np.float64)
            handles.append(L.d2o_tree_import(nodes.ctypes.data_as(_ip), _p(buckets), nodes.shape[0]))
        arr = (_vp * len(handles))(*handles)
        self._tob = np.ascontiguousarray(tree_of_body, dtype=np.int32)
        self._bodies = _d(bodies)
        self.n_bodies = self._tob.size
        self.h = L.d2o_scene_create(len(handles), arr, self.n_bodies, self._tob.ctypes.data_as(_ip), _p(self._bodies))

    def __del__(self):
        if getattr(self, "h", None):
            lib().d2o_scene_free(self.h)
            self.h = None

    def bodies(self):
        o = np.zeros((self.n_bodies, 40))
        lib().d2o_scene_get_bodies(self.h, _p(o))
        return o

    def _result(self, n_pairs, n, per_pair):
        off = np.zeros(n_pairs + 1, dtype=np.int32)
        ints = np.zeros((n, 8), dtype=np.int32)
        reals = np.zeros((n, 11))
        pp = np.zeros(per_pair)
        lib().d2o_scene_result(self.h, off.ctypes.data_as(_ip), ints.ctypes.data_as(_ip), _p(reals), _p(pp))
        return off, ints, reals, pp

    def narrow(self, pairs, filter=True):
        pairs = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
        n = lib().d2o_narrow(self.h, pairs.shape[0], pairs.ctypes.data_as(_ip), int(bool(filter)))
        off, ints, reals, _ = self._result(pairs.shape[0], n, pairs.shape[0])
        return off, ints, reals

    def ccd(self, pairs, max_time):
        pairs = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
        mt = _d(max_time)
        n = lib().d2o_ccd(self.h, pairs.shape[0], pairs.ctypes.data_as(_ip), _p(mt))
        off, ints, reals, pp = self._result(pairs.shape[0], n, 2 * pairs.shape[0])
        return off, ints, reals, pp.reshape(-1, 2)

    def exhaustive_leaves(self, a, b, continuous=False, max_time=0.0):
        """Every leaf pair of bodies a, b through the discrete (or, with
        continuous, the connected CCD) bucket comparison: (ints, reals)."""
        n = lib().d2o_exhaustive_leaves(self.h, int(a), int(b), 1 if continuous else 0, float(max_time))
        _, ints, reals, _ = self._result(1, n, 2)
        return ints, reals

    def filter_last(self, continuous):
        """filterContacts (discrete) / filterContinuousContacts on the last result."""
        n = lib().d2o_scene_filter_result(self.h, 1 if continuous else 0)
        off, ints, reals, _ = self._result(1, n, 0)
        return ints, reals

    def cluster_step(self, pairs, dt):
        pairs = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
        return lib().d2o_cluster_step(self.h, pairs.shape[0], pairs.ctypes.data_as(_ip), float(dt))

    def set_bodies(self, records):
        """Replace every body's record (the warm store stays)."""
        self._bodies = _d(records, 40 * self.n_bodies)
        lib().d2o_scene_set_bodies(self.h, _p(self._bodies))

    def solver_props(self, mass, friction):
        """Per-body mass and friction coefficient for the solver."""
        self._mass, self._fric = _d(mass, self.n_bodies), _d(friction, self.n_bodies)
        lib().d2o_solver_props(self.h, _p(self._mass), _p(self._fric))

    def solve_cluster(self, bodies, pairs, t, g=(0.0, 0.0, -9.8)):
        """SequentialImpulses::solve for one cluster (its bodies, its pairs in
        order) at step t: (solved, per-hit [n, 11]: normal, impulse, impulse
        offset, friction impulse, tangential velocity, in gathered order)."""
        bodies = np.ascontiguousarray(bodies, dtype=np.int32).ravel()
        pairs = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
        gv = _d(g, 3)
        nh = ctypes.c_int(0)
        ok = lib().d2o_solve_cluster(self.h, bodies.size, bodies.ctypes.data_as(_ip), pairs.shape[0],
                                     pairs.ctypes.data_as(_ip), float(t), _p(gv), ctypes.byref(nh))
        out = np.zeros((nh.value, 11))
        if ok:
            lib().d2o_solve_results(self.h, _p(out))
        return bool(ok), out

    def warm(self, a, b):
        """The warm store's impulse for bodies (a, b), or None."""
        v = np.zeros(1)
        return float(v[0]) if lib().d2o_warm_get(self.h, int(a), int(b), _p(v)) else None

    def commit(self, step):
        """current <- future, future <- extrapolate(step) for dynamic bodies."""
        st = _d(step, self.n_bodies)
        lib().d2o_commit(self.h, _p(st))

    def time_pass(self, pairs, max_time, threads=1, repeats=5):
        pairs = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
        mt = _d(max_time)
        cc = np.zeros(1)
        t = lib().d2o_time_pass(self.h, pairs.shape[0], pairs.ctypes.data_as(_ip), _p(mt), threads, repeats, _p(cc))
        return t, int(cc[0])


def filter_order(reals, margin=0.0):
    """filterContacts on contact reals (n x 11): (the inputs' order after the
    in-place std::sort, the kept inputs in that order)."""
    r = _d(reals)
    n = r.size // 11
    order = np.zeros(max(n, 1), dtype=np.int32)
    kept = np.zeros(max(n, 1), dtype=np.int32)
    k = lib().d2o_filter_order(_p(r), n, float(margin), order.ctypes.data_as(_ip), kept.ctypes.data_as(_ip))
    return order[:n], kept[:k]


def step_scale(reals, body_a, body_b, max_time):
    """TOI -> step rule on explicit contact reals (n x 11)."""
    r = _d(reals)
    return lib().d2o_step_scale(_p(r), r.size // 11, _p(_d(body_a, 40)), _p(_d(body_b, 40)), float(max_time))


# --- broad phase and clusters ------------------------------------------------
def _opt(a, dtype, ptype):
    if a is None:
        return None, None
    a = np.ascontiguousarray(np.asarray(a, dtype=dtype).ravel())
    return a, a.ctypes.data_as(ptype)


def mesh_bounds(V):
    """MeshData::initBoudingData: (float32 AABB [lo xyz, hi xyz], radius)."""
    V = _d(V)
    box = np.zeros(6, dtype=np.float32)
    r = ctypes.c_double(0)
    lib().d2o_mesh_bounds(_p(V), V.size // 3, box.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), ctypes.byref(r))
    return box, r.value


def broad_phase(records, aabb6, radius, mode, t=0.0, last=None, groups=None):
    """broadPhaseEmbree by brute force: mode 0 = (particles, t), 1 = (particles).
    aabb6/radius are per body.  Returns int32 pairs (a < b, sorted); raises
    RuntimeError when State::interpolate throws."""
    records = _d(records)
    n = records.size // 40
    last, lp = _opt(last, np.float64, _dp)
    box, bp = _opt(aabb6, np.float32, ctypes.POINTER(ctypes.c_float))
    rad, rp = _opt(radius, np.float64, _dp)
    grp, gp = _opt(groups, np.int32, _ip)
    cap = 1 << 16
    while True:
        out = np.zeros(2 * cap, dtype=np.int32)
        k = lib().d2o_broad_phase(n, _p(records), lp, bp, rp, gp, int(mode), float(t), cap, out.ctypes.data_as(_ip))
        if k < 0:
            raise RuntimeError("Invalid args for interpolation")
        if k <= cap:
            return out[:2 * k].reshape(-1, 2)
        cap = k


def clusters(records, pairs, last=None):
    """separateCollisionClusters: dict with n, of_body, of_pair, first,
    min_time, step and the rolled-back current states (n x 14)."""
    records = _d(records)
    n = records.size // 40
    pairs = np.ascontiguousarray(np.asarray(pairs, dtype=np.int32).reshape(-1, 2))
    last, lp = _opt(last, np.float64, _dp)
    ob = np.zeros(n, dtype=np.int32)
    op = np.zeros(max(pairs.shape[0], 1), dtype=np.int32)
    first = np.zeros(n, dtype=np.int32)
    mt = np.zeros(n)
    st = np.zeros(n)
    cur = np.zeros(14 * n)
    k = lib().d2o_clusters(n, _p(records), lp, pairs.shape[0], pairs.ctypes.data_as(_ip), ob.ctypes.data_as(_ip),
                           op.ctypes.data_as(_ip), first.ctypes.data_as(_ip), _p(mt), _p(st), _p(cur))
    if k < 0:
        raise RuntimeError("Invalid args for interpolation")
    return {"n": k, "of_body": ob, "of_pair": op[:pairs.shape[0]], "first": first[:k], "min_time": mt[:k],
            "step": st[:k], "current": cur.reshape(n, 14)}


def interpolate(cur14, last14, time):
    """State::interpolate; raises RuntimeError when the reference throws."""
    out = np.zeros(14)
    if lib().d2o_interpolate(_p(_d(cur14)), _p(_d(last14)), float(time), _p(out)) < 0:
        raise RuntimeError("Invalid args for interpolation")
    return out
