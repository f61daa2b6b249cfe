"""The device solver's code (d2_solver.hpp, d2_sort.hpp), built for the host
with g++, against the oracle's restatement of SequentialImpulses::solve: the
same clusters, hits gathered as the device gathers them, bit-for-bit equal
future states and per-contact results, over two consecutive steps (the second
starts from the warm store the first left)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components

import d2o
from gpu_cases import c1_scene, dense_scene, hull_scene, with_future
from delta2_amd import scenes
from test_oracle_pins import _two_leaf_scene, record

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
G = np.array([0.0, 0.0, -9.8])
_ip = ctypes.POINTER(ctypes.c_int32)
_dp = ctypes.POINTER(ctypes.c_double)
_up = ctypes.POINTER(ctypes.c_uint64)


@pytest.fixture(scope="module")
def host(tmp_path_factory):
    so = str(tmp_path_factory.mktemp("solver") / "solver_host.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-shared", "-fPIC",
                    os.path.join(ROOT, "tests", "solver_host_check.cpp"), "-o", so], check=True)
    L = ctypes.CDLL(so)
    L.solver_host.argtypes = [ctypes.c_int, _dp, _dp, _dp, _ip, _ip, ctypes.c_int, _ip, _dp, _dp, _dp, ctypes.c_double,
                              _dp, _up, _dp, ctypes.c_uint64, _up, _dp]
    L.solver_host.restype = ctypes.c_int
    return L


class HostSolver:
    """The host build with a warm table that persists across calls."""

    def __init__(self, L, cap=1 << 12):
        self.L = L
        self.keys = np.full(cap, np.iinfo(np.uint64).max, dtype=np.uint64)
        self.vals = np.zeros(cap)
        self.used = np.zeros(1, dtype=np.uint64)

    def solve(self, records, mass, fric, pairs, reals, t):
        """One cluster: pairs in order, reals the filtered contacts of each
        pair (a list of [n, 11] arrays).  Returns (ok, out[n, 11]); records
        are updated in place."""
        part = []
        loc = {}

        def local(b):
            if b not in loc:
                loc[b] = len(part)
                part.append(b)
            return loc[b]

        hA, hB, heps, hl, hpc = [], [], [], [], []
        for (a, b), r in zip(pairs, reals):
            la, lb = local(int(a)), local(int(b))
            for row in r:
                hA.append(row[0:3]), hB.append(row[3:6]), heps.append(row[6] + row[7]), hl.append((la, lb))
                hpc.append(len(r))
        nh = len(hA)
        hA = np.ascontiguousarray(np.array(hA, dtype=float).reshape(-1, 3))
        hB = np.ascontiguousarray(np.array(hB, dtype=float).reshape(-1, 3))
        heps = np.ascontiguousarray(np.array(heps, dtype=float))
        hl = np.ascontiguousarray(np.array(hl, dtype=np.int32).reshape(-1, 2))
        hpc = np.ascontiguousarray(np.array(hpc, dtype=np.int32))
        pb = np.ascontiguousarray(np.array(part, dtype=np.int32))
        out = np.zeros((max(nh, 1), 11))
        gv = np.ascontiguousarray(G)
        ok = self.L.solver_host(nh, hA.ctypes.data_as(_dp), hB.ctypes.data_as(_dp), heps.ctypes.data_as(_dp),
                                hl.ctypes.data_as(_ip), hpc.ctypes.data_as(_ip), pb.size, pb.ctypes.data_as(_ip), records.ctypes.data_as(_dp),
                                mass.ctypes.data_as(_dp), fric.ctypes.data_as(_dp), float(t), gv.ctypes.data_as(_dp),
                                self.keys.ctypes.data_as(_up), self.vals.ctypes.data_as(_dp), self.keys.size,
                                self.used.ctypes.data_as(_up), out.ctypes.data_as(_dp))
        return bool(ok), out[:nh]


def clusters(pairs, n):
    g = coo_matrix((np.ones(len(pairs)), (pairs[:, 0], pairs[:, 1])), shape=(n, n))
    _, lab = connected_components(g, directed=False)
    return [np.nonzero(lab[pairs[:, 0]] == k)[0] for k in np.unique(lab[pairs[:, 0]])]


def run_both(L, trees, tob, records, pairs, mass, fric, t, steps=2):
    """Solve every cluster with the oracle and with the host build of the
    device code, `steps` times with a commit in between; compare bits."""
    osc = d2o.Scene(trees, tob, records)
    osc.solver_props(mass, fric)
    hs = HostSolver(L)
    rec_h = np.ascontiguousarray(records.copy())
    solved = failed = contacts = 0
    for step in range(steps):
        view = d2o.Scene(trees, tob, rec_h)  # the host side's contacts, at its own future states
        off, _, reals = view.narrow(pairs, filter=True)
        for idx in clusters(pairs, records.shape[0]):
            cp = pairs[idx]
            bodies = np.unique(cp)
            ok_o, out_o = osc.solve_cluster(bodies, cp, t)
            ok_h, out_h = hs.solve(rec_h, mass, fric, cp, [reals[off[p]:off[p + 1]] for p in idx], t)
            assert ok_o == ok_h, (step, idx)
            if ok_o:
                assert np.array_equal(out_h, out_o), (step, idx, np.abs(out_h - out_o).max())
                solved += 1
                contacts += out_o.shape[0]
            else:
                failed += 1
        assert np.array_equal(rec_h[:, 14:28], osc.bodies()[:, 14:28]), step
        for (a, b) in {tuple(sorted(p)) for p in pairs.tolist()}:  # the warm stores agree
            w = osc.warm(a, b)
            k = (min(a, b) << 32) | max(a, b)
            hit = np.nonzero(hs.keys == np.uint64(k))[0]
            assert (w is None) == (hit.size == 0) and (w is None or hs.vals[hit[0]] == w)
        # commit: current <- future, future <- extrapolate (both sides alike)
        step_t = np.full(records.shape[0], t)
        osc.commit(step_t)
        rec_h[:, 0:14] = osc.bodies()[:, 0:14]
        rec_h[:, 14:28] = osc.bodies()[:, 14:28]
    return solved, failed, contacts


def test_pile_cases(host):
    for z, vel, x, fr in [(1.103, (0, 0, -1.0), 0.0, 0.0), (1.103, (0, 0, -1.0), -0.3, 0.0),
                          (1.103, (1.0, 0, -1.0), 0.0, 0.5), (1.115, (0, 0, 1.0), 0.0, 0.5)]:
        sc = _two_leaf_scene([record(trans=(0, 0, 1.0), static=True), record(trans=(x, 0, z), vel=vel)])
        from test_oracle_kat import PLANE_F, PLANE_V, PYRAMID_F, PYRAMID_V
        rng = d2o.pcg_seed(42, 54)
        trees = [d2o.build_tree(PYRAMID_V, PYRAMID_F, rng), d2o.build_tree(PLANE_V, PLANE_F, rng)]
        rec = sc.bodies()
        full = np.stack([record(trans=(0, 0, 1.0), static=True), record(trans=(x, 0, z), vel=vel)])
        full[:, 0:28] = rec[:, 0:28]
        s, f, n = run_both(host, trees, [0, 1], full, np.array([[0, 1]], dtype=np.int32), np.ones(2),
                           np.full(2, fr), 0.01)
        assert s == 2 and n >= 1


# t = 1e-4 lets the scenes' touching contacts through (the solver refuses
# |A - B| < 1e-6 only for t > 1e-4); at the frame step every cluster of the
# 100-body scene with a touching contact is refused by both
@pytest.mark.parametrize("make,t", [(dense_scene, 1e-4), (hull_scene, 1e-4), (c1_scene, 1e-4), (c1_scene, scenes.DT)])
def test_scene_clusters(host, make, t):
    s = make()
    rec = np.ascontiguousarray(with_future(s.records, scenes.DT))
    rng = np.random.default_rng(5)
    mass = rng.uniform(0.5, 2.0, rec.shape[0])
    fric = rng.uniform(0.0, 0.6, rec.shape[0])
    solved, failed, contacts = run_both(host, s.trees, s.tree_of_body, rec, s.pairs, mass, fric, t)
    print(f"[solver cpu] {make.__name__}: {solved} clusters solved, {failed} refused, {contacts} contacts")
    assert (solved > 0 and contacts > 0) if t < 1e-3 else failed > 0
