"""GPU: the device solver (d2_solve, sequential impulses over every cluster of
a filtered step) against the oracle's restatement of
SequentialImpulses::solve, over two consecutive steps (step, solve, commit,
step, solve: the second solve starts from the warm store of the first).

The oracle solves each cluster (its pairs in index order, its bodies) at the
cluster step the device chose, from the body states the device's step left;
bodies in no pair are clusters without contacts.  Compared: refusals exactly,
per-contact normals, impulses, friction impulses and tangential velocities,
and the new future states to 1e-9 relative (the device's and the C
library's sin/cos may differ in the last bit, and the solver iterates on the
turned poses).  Scenes: a plane landing on a pyramid (straight, off centre,
sliding with friction, separating) and a 4 x 4 grid of spheres settling on a
static floor with 3 mm gaps (clusters as the reference forms them: static
bodies join every cluster they touch)."""
import numpy as np
import pytest

import d2o
import delta2_amd as d2
from delta2_amd import scenes
from test_oracle_kat import PLANE_F, PLANE_V, PYRAMID_F, PYRAMID_V
from test_oracle_pins import record

pytestmark = pytest.mark.gpu
G = np.array([0.0, 0.0, -9.8])
DT32 = float(np.float32(scenes.DT))


def ref_clusters(pairs, static):
    """separate_clusters.cpp:355-459: components over dynamic-dynamic pairs; a
    pair with a static body belongs to its dynamic body's component."""
    n = static.size
    parent = np.arange(n)

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    for a, b in pairs:
        if not static[a] and not static[b]:
            ra, rb = find(a), find(b)
            if ra != rb:
                parent[max(ra, rb)] = min(ra, rb)
    root = np.array([find(b if static[a] else a) for a, b in pairs])
    used, cl = np.unique(root, return_inverse=True)
    return cl.astype(np.int32), len(used)


def close(a, b, tol=1e-9):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b) <= tol * np.maximum(np.abs(b), 1e-3)


def run_case(label, trees, tob, rec, pairs, mass, fric, steps=2):
    pairs = np.asarray(pairs, dtype=np.int32).reshape(-1, 2)
    static = rec[:, 38] != 0.0
    cl, k = ref_clusters(pairs, static)
    osc = d2o.Scene(trees, tob, rec)
    osc.solver_props(mass, fric)
    totals = {"solved": 0, "refused": 0, "contacts": 0, "bitwise": 0, "values": 0}
    with d2.Context(0, d2.D2_FP64) as ctx:
        ctx.upload_trees(trees)
        ctx.upload_bodies(tob, rec)
        ctx.upload_pairs(pairs)
        ctx.upload_clusters(cl, k)
        ctx.solver_bodies(mass, fric)
        for step in range(steps):
            last = ctx.download_bodies()[:, 39].copy()
            n, csteps = ctx.step_run(scenes.DT, filter=True, want_steps=True)
            before = ctx.download_bodies()
            res = ctx.result(with_pairs=True)
            status, n_failed = ctx.solve(G)
            out = ctx.solve_results()
            after = ctx.download_bodies()
            print(f"[solver {label}] step {step}: {k} clusters, {n} contacts, {n_failed} refused", flush=True)
            assert out.shape[0] == n
            osc.set_bodies(before)
            body_step = np.clip(last, 1e-3, DT32)
            in_pair = np.zeros(rec.shape[0], dtype=bool)
            in_pair[pairs.ravel()] = True
            for j in range(k):
                idx = np.nonzero(cl == j)[0]
                cp = pairs[idx]
                ok, o_out = osc.solve_cluster(np.unique(cp), cp, csteps[j], G)
                assert bool(status[j]) == ok, (label, step, j)
                body_step[np.unique(cp)] = csteps[j]
                if not ok:
                    totals["refused"] += 1
                    continue
                rows = np.concatenate([np.arange(res.offsets[p], res.offsets[p + 1]) for p in idx]) if idx.size else []
                got = out[np.asarray(rows, dtype=int)]
                assert got.shape == o_out.shape
                assert close(got, o_out).all(), (label, step, j, np.abs(got - o_out).max())
                totals["solved"] += 1
                totals["contacts"] += got.shape[0]
                totals["bitwise"] += int((got == o_out).sum())
                totals["values"] += got.size
            for b in np.nonzero(~in_pair & ~static)[0]:  # free bodies: gravity only
                ok, _ = osc.solve_cluster([b], np.zeros((0, 2)), body_step[b], G)
                assert ok
            want = osc.bodies()
            assert close(after[:, 14:28], want[:, 14:28]).all(), (label, step, np.abs(after[:, 14:28] - want[:, 14:28]).max())
            assert np.array_equal(after[static], before[static])
            ctx.commit_step()
            com = ctx.download_bodies()
            assert np.array_equal(com[~static, 0:14], after[~static, 14:28])  # current <- future
            assert np.array_equal(com[~static, 39], body_step[~static])
    print(f"[solver {label}] {totals}", flush=True)
    return totals


def pyramid_plane_trees():
    rng = d2o.pcg_seed(42, 54)
    return [d2o.build_tree(PYRAMID_V, PYRAMID_F, rng), d2o.build_tree(PLANE_V, PLANE_F, rng)]


# descending at 0.1 m/s the plane stays clear of the apex over a frame (the
# contact is pushed, not refused); the last case separates
@pytest.mark.parametrize("z,vel,x,fr", [(1.103, (0, 0, -0.1), 0.0, 0.0), (1.103, (0, 0, -0.1), -0.3, 0.0),
                                        (1.103, (0.5, 0, -0.1), 0.0, 0.5), (1.115, (0, 0, 1.0), 0.0, 0.5)])
def test_plane_on_pyramid(z, vel, x, fr):
    rec = np.stack([record(trans=(0, 0, 1.0), static=True, last=scenes.DT),
                    record(trans=(x, 0, z), vel=vel, last=scenes.DT)])
    t = run_case(f"pyramid z={z} x={x}", pyramid_plane_trees(), [0, 1], rec, [[0, 1]], np.ones(2), np.full(2, fr))
    assert t["solved"] >= 1


def grid_scene(n=4, gap=0.003, r=0.5):
    """A static 20 m floor (two triangles) and n x n spheres (subdivision 1)
    above it, `gap` apart from the floor and from their neighbours, falling
    slowly with a little sideways motion."""
    rng = np.random.default_rng(9)
    FV = np.array([[-10, -10, 0], [10, -10, 0], [10, 10, 0], [-10, 10, 0]], dtype=float)
    FF = np.array([[0, 1, 2], [0, 2, 3]], dtype=np.int32)
    V, F = scenes.icosphere(1, r)
    vol, _, inertia = scenes.mass_properties(V, F)
    stream = d2.pcg_seed(42, 54)
    trees = [scenes.tree_build(FV, FF, stream), scenes.tree_build(V, F, stream)]
    recs = [record(trans=(0, 0, 0), static=True, last=scenes.DT)]
    low = -float(V[:, 2].min())  # the mesh's lowest point below its centre (identity orientation)
    sx, sy = 2 * float(V[:, 0].max()) + gap, 2 * float(V[:, 1].max()) + gap
    for i in range(n):
        for j in range(n):
            pos = ((i - n / 2) * sx, (j - n / 2) * sy, low + gap)
            v = np.array([rng.normal(0, 0.05), rng.normal(0, 0.05), -0.05])
            recs.append(record(trans=pos, vel=v, inv_inertia=np.linalg.inv(inertia), last=scenes.DT))
    rec = np.stack(recs)
    tob = [0] + [1] * (n * n)
    radii = np.array([15.0] + [r] * (n * n))
    pairs = scenes.broad_phase(rec, radii)
    mass = np.array([1.0] + [vol] * (n * n))
    return trees, tob, rec, pairs, mass


def test_sphere_grid_on_floor():
    trees, tob, rec, pairs, mass = grid_scene()
    fric = np.full(rec.shape[0], 0.4)
    t = run_case("grid", trees, tob, rec, pairs, mass, fric)
    assert t["solved"] >= 1 and t["contacts"] > 16
