"""GPU: the benchmark workloads at their stated sizes.

- the 4096-body mixed workload (subdivision-3 icospheres and ellipsoid hulls,
  2.9 M triangles), the batch of 256 independent scenes of 16-512 bodies, the
  4096-body held-out scene of the learned gate, and the 65536-body funnel
  pile (84 M triangles): one composite step (CCD, step rule, cluster steps,
  projection, discrete narrow phase + filter) in the FP32 form with the
  learned gate (committed weights) and subtree pruning equals the plain FP64
  step without pruning bit for bit over every pair, and a seeded sample of 64
  pairs equals the oracle: the per-pair CCD values (relative 1e-9) and the
  discrete contacts after the filter on the future states the device used
  (index tuples exact, reals relative 1e-9).
The oracle gets only the sampled pairs' bodies and trees.
"""
import numpy as np
import pytest

import d2o
import delta2_amd as d2
from delta2_amd import learned, scenes

pytestmark = pytest.mark.gpu
TOL = 1e-9
SAMPLE = 64


def close(a, b, tol=TOL):
    """Relative bound |a - b| <= tol |b| (exact where b is zero)."""
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b) <= tol * np.abs(b)


def clusters_of(pairs, n_bodies):
    """Connected components of the pair graph (the caller's clustering)."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components
    g = coo_matrix((np.ones(pairs.shape[0]), (pairs[:, 0], pairs[:, 1])), shape=(n_bodies, n_bodies))
    _, label = connected_components(g, directed=False)
    used, cl = np.unique(label[pairs[:, 0]], return_inverse=True)
    return cl.astype(np.int32), len(used)


def one_step(s, rec, pairs, cl, k, precision, weights=None, prune=True):
    """One d2_step_run on a fresh context: (result, cluster steps, bodies after)."""
    with d2.Context(0, precision) as c:
        c.set_pruning(prune)
        if weights is not None:
            c.mlp_set_params(weights)
            c.set_learned(True)
        c.upload_trees(s.trees)
        c.upload_bodies(s.tree_of_body, rec)
        c.upload_pairs(pairs)
        c.upload_clusters(cl, k)
        n, steps = c.step_run(scenes.DT, filter=True, want_steps=True)
        r = c.result(with_pairs=True)
        out = (r.offsets.copy(), r.ints.copy(), r.reals.copy(), r.per_pair.copy())
        gate = c.gate_counts() if weights is not None else None
        return out, steps.copy(), c.download_bodies(), n, gate


def oracle_of(s, records, pairs):
    """The oracle over the bodies of `pairs` only: (scene, local pairs)."""
    bodies, inv = np.unique(pairs, return_inverse=True)
    tids, tinv = np.unique(s.tree_of_body[bodies], return_inverse=True)
    sc = d2o.Scene([s.trees[int(t)] for t in tids], tinv.astype(np.int32), records[bodies])
    return sc, inv.reshape(-1, 2).astype(np.int32)


def check_config(s, label, seed):
    weights = learned.load_weights()
    assert weights is not None, "the committed gate weights are missing"
    rec = s.records
    pairs = s.pairs
    cl, k = clusters_of(pairs, s.n_bodies)
    print(f"[{label}] {s.n_bodies} bodies, {pairs.shape[0]} pairs, {k} clusters", flush=True)
    plain, steps64, bodies64, n64, _ = one_step(s, rec, pairs, cl, k, d2.D2_FP64, prune=False)
    print(f"[{label}] plain FP64 step: {n64} contacts", flush=True)
    fast, steps32, bodies32, n32, gate = one_step(s, rec, pairs, cl, k, d2.D2_FP32, weights)
    print(f"[{label}] FP32 + gate step: {n32} contacts, gate {gate}", flush=True)
    # the fast form (FP32 culls, learned gate, subtree pruning) equals the
    # plain one (FP64, every node pair of the reference's traversal) bit for
    # bit, over every pair
    for name, a, b in zip(("offsets", "ints", "reals", "per_pair"), fast, plain):
        assert np.array_equal(a, b), f"{label}: FP32 + gate {name} differ from plain FP64"
    assert np.array_equal(steps32, steps64) and np.array_equal(bodies32, bodies64)
    # a seeded sample against the oracle
    rng = np.random.default_rng(seed)
    sel = np.sort(rng.choice(pairs.shape[0], size=min(SAMPLE, pairs.shape[0]), replace=False))
    off, ints, reals, per_pair = plain
    mt = np.minimum(rec[pairs[sel, 0], 39], rec[pairs[sel, 1], 39])
    osc, lp = oracle_of(s, rec, pairs[sel])
    _, _, _, o_pp = osc.ccd(lp, mt)
    assert close(per_pair[sel], o_pp).all(), f"{label}: per-pair CCD values differ from the oracle"
    osc, lp = oracle_of(s, bodies64, pairs[sel])
    o_off, o_ints, o_reals = osc.narrow(lp, filter=True)
    got_i, got_r = [], []
    for j, p in enumerate(sel):
        gi = ints[off[p]:off[p + 1]].copy()
        gi[:, 0] = j  # the pair column names the sampled pair
        got_i.append(gi)
        got_r.append(reals[off[p]:off[p + 1]])
    got_i, got_r = np.concatenate(got_i), np.concatenate(got_r)
    assert np.array_equal(np.diff(o_off), off[sel + 1] - off[sel]), f"{label}: contacts per pair differ"
    assert np.array_equal(got_i, o_ints), f"{label}: contact index tuples differ"
    assert close(got_r, o_reals).all(), f"{label}: contact reals differ"
    print(f"[{label}] sample of {sel.size} pairs: {o_ints.shape[0]} contacts, "
          f"{int((o_pp[:, 0] < 1.0).sum())} pairs with a scaled step", flush=True)


@pytest.mark.timeout(300)
def test_mixed_4096_body_workload():
    check_config(scenes.scene_c2(), "c2 4096", seed=2)


@pytest.mark.timeout(300)
def test_batched_256_scenes():
    check_config(scenes.scene_c3(), "c3 256 scenes", seed=3)


@pytest.mark.timeout(300)
def test_held_out_4096_body_scene():
    check_config(scenes.scene_c4(), "c4 4096", seed=4)


@pytest.mark.timeout(900)
def test_funnel_pile_65536_bodies():
    s = scenes.scene_c5(progress=lambda m: print(f"[c5] {m}", flush=True))
    check_config(s, "c5 65536", seed=5)
