"""GPU: the composite step, graph replays and launch shapes.

- d2_step_run (CCD + step rule + cluster steps + projection + discrete narrow
  phase + filter, all on the device) against the oracle: cluster steps to
  1e-12 relative, future states to 1e-14 (device and glibc sin/cos may differ
  in the last bit), and the contacts bit-for-bit against the oracle's narrow
  phase on the states the device used.
- Graph mode: every replay equals the eager run, with eager work and
  synchronisations between replays.
- Launch shapes (d2_tune) change nothing: one block per chunked kernel makes
  the in-kernel scans loop over many tiles.
"""
import numpy as np
import pytest

import delta2_amd as d2
from gpu_cases import c1_scene, dense_scene, oracle_scene, sparse_scene, with_future
from delta2_amd import scenes

pytestmark = pytest.mark.gpu
TOL = 1e-9


def close(a, b, tol=TOL):
    """Relative bound |a - b| <= tol |b| (exact where b is zero)."""
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b) <= tol * np.abs(b)


def same_result(a, b, label):
    assert np.array_equal(a.offsets, b.offsets), f"{label}: offsets differ"
    assert np.array_equal(a.ints, b.ints), f"{label}: index tuples differ"
    assert np.array_equal(a.reals, b.reals), f"{label}: reals differ"
    if a.per_pair is not None and b.per_pair is not None:
        assert np.array_equal(a.per_pair, b.per_pair), f"{label}: per-pair values differ"


def load(ctx, s, records):
    ctx.upload_trees(s.trees)
    ctx.upload_bodies(s.tree_of_body, records)


def clusters_of(pairs, n_bodies):
    """Connected components of the pair graph (separate_clusters.cpp:355-459 is
    the caller's job): every body belongs to exactly one cluster."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components
    g = coo_matrix((np.ones(pairs.shape[0]), (pairs[:, 0], pairs[:, 1])), shape=(n_bodies, n_bodies))
    _, label = connected_components(g, directed=False)
    used, cl = np.unique(label[pairs[:, 0]], return_inverse=True)
    return cl.astype(np.int32), len(used)


@pytest.fixture(scope="module")
def ctx():
    c = d2.Context(0, d2.D2_FP64)
    yield c
    c.close()


@pytest.mark.parametrize("name,make", [("c1", c1_scene), ("dense", dense_scene), ("c2_256", sparse_scene)])
def test_step_matches_oracle(ctx, name, make):
    s = make()
    print(f"[step] scene={name} pairs={s.pairs.shape[0]}", flush=True)
    rec = s.records.copy()
    rec[::7, 39] = scenes.DT / 4  # uneven last steps: max_time = min of the two bodies
    load(ctx, s, rec)
    pairs = s.pairs
    cl, k = clusters_of(pairs, s.n_bodies)
    ctx.upload_pairs(pairs)
    ctx.upload_clusters(cl, k)
    n, steps = ctx.step_run(scenes.DT, filter=True, want_steps=True)
    print(f"[step] clusters={k} contacts={n}", flush=True)
    # cluster steps: the oracle runs fineCollisionClustersWithTimeStepSelection per cluster
    osc = oracle_scene(s, rec)
    want = np.array([osc.cluster_step(pairs[cl == j], np.float32(scenes.DT)) for j in range(k)])
    assert np.allclose(steps, want, rtol=1e-12, atol=0), (steps, want)
    # future states: cluster step for bodies in pairs; a body in no pair is a
    # cluster of its own and steps with its clamped last step
    body_step = np.clip(rec[:, 39], 1e-3, float(np.float32(scenes.DT)))
    body_step[pairs[:, 0]] = steps[cl]
    body_step[pairs[:, 1]] = steps[cl]
    got_bodies = ctx.download_bodies()
    want_bodies = with_future(rec, body_step)
    assert np.allclose(got_bodies[:, 14:28], want_bodies[:, 14:28], rtol=0, atol=1e-14)
    # contacts: the oracle's narrow phase + filter on the states the device used
    gpu = ctx.result(with_pairs=True)
    assert gpu.n == n
    o_off, o_ints, o_reals = oracle_scene(s, got_bodies).narrow(pairs, filter=True)
    assert np.array_equal(gpu.offsets, o_off)
    assert np.array_equal(gpu.ints, o_ints)
    assert close(gpu.reals, o_reals).all()
    # the per-pair CCD values of the step equal a plain CCD run
    mt = np.minimum(rec[pairs[:, 0], 39], rec[pairs[:, 1], 39])
    _, _, _, o_pp = oracle_scene(s, rec).ccd(pairs, mt)
    assert close(gpu.per_pair, o_pp).all()


@pytest.mark.parametrize("precision", [d2.D2_FP64, d2.D2_FP32])
def test_graph_replays_equal_eager(precision):
    s = dense_scene()
    print(f"[graph] precision={precision}", flush=True)
    c = d2.Context(0, precision)
    try:
        rec = with_future(s.records, scenes.DT)
        load(c, s, rec)
        pairs = s.pairs
        mt = np.minimum(rec[pairs[:, 0], 39], rec[pairs[:, 1], 39])
        cl, k = clusters_of(pairs, s.n_bodies)
        # eager references
        c.upload_pairs(pairs, mt)
        c.narrow_run(True)
        e_narrow = c.result()
        c.ccd_run()
        e_ccd = c.result(with_pairs=True)
        c.upload_clusters(cl, k)
        _, e_steps = c.step_run(scenes.DT, True, want_steps=True)
        e_step = c.result(with_pairs=True)
        e_bodies = c.download_bodies()
        c.set_graph(True)
        for rep in range(3):
            load(c, s, rec)  # eager work (uploads, synchronisation) between replays
            c.upload_pairs(pairs, mt)
            c.upload_clusters(cl, k)
            c.narrow_run(True)
            same_result(c.result(), e_narrow, f"narrow replay {rep}")
            c.ccd_run()
            same_result(c.result(with_pairs=True), e_ccd, f"ccd replay {rep}")
            c.project_future(np.full(s.n_bodies, scenes.DT))  # an eager kernel between replays
            _, st = c.step_run(scenes.DT, True, want_steps=True)
            assert np.array_equal(st, e_steps), f"cluster steps, replay {rep}"
            same_result(c.result(with_pairs=True), e_step, f"step replay {rep}")
            assert np.array_equal(c.download_bodies(), e_bodies)
        c.set_graph(False)
    finally:
        c.close()


@pytest.mark.parametrize("shape", [(1, 1, 1), (3, 7, 5), (4096, 16384, 16384)])
def test_launch_shapes_change_nothing(ctx, shape):
    s = dense_scene()
    print(f"[tune] shape={shape}", flush=True)
    rec = with_future(s.records, scenes.DT)
    load(ctx, s, rec)
    pairs = s.pairs
    mt = np.full(pairs.shape[0], 4 * scenes.DT)
    default = ctx.get_tune()
    try:
        ctx.upload_pairs(pairs, mt)
        ctx.narrow_run(True)
        ref_n = ctx.result()
        ctx.ccd_run()
        ref_c = ctx.result(with_pairs=True)
        ctx.set_tune(*shape)
        ctx.narrow_run(True)
        same_result(ctx.result(), ref_n, f"narrow {shape}")
        ctx.ccd_run()
        same_result(ctx.result(with_pairs=True), ref_c, f"ccd {shape}")
    finally:
        ctx.set_tune(*default)


def test_step_needs_clusters(ctx):
    s = c1_scene()
    load(ctx, s, s.records)
    ctx.upload_pairs(s.pairs)  # uploading pairs clears the clusters
    with pytest.raises(d2.D2Error):
        ctx.step_run(scenes.DT)
    with pytest.raises(d2.D2Error):
        ctx.upload_clusters(np.full(s.pairs.shape[0], 5, dtype=np.int32), 2)


def test_graph_continues_past_captured_levels():
    """A graph captured on a one-level pair list, replayed on a deep one: the
    host continues past the captured levels and the result equals eager."""
    s = dense_scene()
    rec = with_future(s.records, scenes.DT)
    pos = rec[:, 4:7]
    d = np.linalg.norm(pos[:, None, :] - pos[None, :, :], axis=-1)
    i, j = np.unravel_index(np.argmax(d), d.shape)
    far = np.array([[min(i, j), max(i, j)]], dtype=np.int32)
    pairs = s.pairs
    mt = np.minimum(rec[pairs[:, 0], 39], rec[pairs[:, 1], 39])
    eager = d2.Context(0, d2.D2_FP64)
    graph = d2.Context(0, d2.D2_FP64)
    try:
        for c in (eager, graph):
            load(c, s, rec)
        eager.upload_pairs(pairs, mt)
        eager.narrow_run(True)
        e_n = eager.result()
        eager.ccd_run()
        e_c = eager.result(with_pairs=True)
        graph.set_graph(True)
        for _ in range(2):  # observe (eager), then capture with one level per phase
            graph.upload_pairs(far, np.full(1, scenes.DT))
            assert graph.narrow_run(True) == 0
            assert graph.ccd_run() == 0
        for rep in range(2):  # deeper lists: continuation, then a recapture with more levels
            print(f"[graph-continue] rep={rep}", flush=True)
            graph.upload_pairs(pairs, mt)
            graph.narrow_run(True)
            same_result(graph.result(), e_n, f"narrow continuation {rep}")
            graph.ccd_run()
            same_result(graph.result(with_pairs=True), e_c, f"ccd continuation {rep}")
    finally:
        eager.close()
        graph.close()


def test_cluster_step_static_first_pair():
    """separate_clusters.cpp:589-590: the second loop starts from step[a] of a
    cluster's first pair; with a static body a that is its unchanged last
    step, and the pair's own scaled step only counts when its dynamic body is
    read later (a later pair's id) or lowered again (a later pair with
    contacts).  A plane (body 1) sinks through a static pyramid's apex
    (body 0) within a 0.01 s interval: scale 0.375; body 2 is a far plane."""
    from test_oracle_kat import PLANE_F, PLANE_V, PYRAMID_F, PYRAMID_V
    from test_oracle_pins import record
    rng = d2.pcg_seed(42, 54)
    trees = [d2.tree_build(PYRAMID_V, PYRAMID_F, rng), d2.tree_build(PLANE_V, PLANE_F, rng)]
    import d2o
    rng = d2o.pcg_seed(42, 54)
    otrees = [d2o.build_tree(PYRAMID_V, PYRAMID_F, rng), d2o.build_tree(PLANE_V, PLANE_F, rng)]
    rec = np.stack([record(trans=(0, 0, 1.0), static=True, last=0.01),
                    record(trans=(0, 0, 1.105), vel=(0, 0, -1.0), last=0.01),
                    record(trans=(5.0, 0, 1.105), last=0.01)])
    tob = [0, 1, 1]
    dt = np.float32(1 / 60)
    cases = [
        (np.array([[0, 1]]), None),            # the pair alone: step[a] = 0.01
        (np.array([[0, 1], [1, 2]]), None),    # body 1 is the id of a later pair: 0.01 * 0.375
        (np.array([[0, 2], [0, 1]]), None),    # static first pair without contacts, then the sinking pair
    ]
    osc = d2o.Scene(otrees, tob, rec)
    with d2.Context(0) as c:
        c.upload_trees(trees)
        c.upload_bodies(tob, rec)
        for pairs, _ in cases:
            mt = np.minimum(rec[pairs[:, 0], 39], rec[pairs[:, 1], 39])
            c.upload_pairs(pairs, mt)
            c.ccd_run()
            got = c.cluster_steps(np.zeros(pairs.shape[0], np.int32), 1, dt)
            want = osc.cluster_step(pairs, dt)
            print(f"[static first] pairs={pairs.tolist()} gpu={got[0]} oracle={want}", flush=True)
            assert got[0] == want, (pairs.tolist(), got, want)
    # the three cases really differ: the quirk is exercised
    steps = [osc.cluster_step(p, dt) for p, _ in cases]
    assert steps[0] == 0.01 and abs(steps[1] - 0.00375) < 1e-15
