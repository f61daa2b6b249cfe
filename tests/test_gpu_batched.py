"""GPU: many independent scenes in one step (the batched workload at small size).

8 scenes of 16-64 bodies (C1 meshes, one gravity direction each) go through
one d2_step_run.  Checked: the batched result restricted to a scene equals
that scene run alone; cluster steps and contacts equal the oracle; the FP32
cull with graph replay gives exactly the FP64 result; the autotuner leaves the
shape it reports set, and the result does not change under it.  A larger
batch (12 scenes of 256-512 bodies, 7.8 k pairs) takes the post-processing's
separate scan launch; its scenes alone (below 4096 pairs) take the scan inside
the scatter kernel, and both give the same contacts and steps.
"""
import functools

import numpy as np
import pytest

import delta2_amd as d2
from gpu_cases import oracle_scene
from delta2_amd import scenes

pytestmark = pytest.mark.gpu
TOL = 1e-9


@functools.lru_cache(maxsize=None)
def small_batch():
    return scenes.scene_c3(seed=1103, n_scenes=8, lo=16, hi=64, threads=4)


def close(a, b, tol=TOL):
    """Relative bound |a - b| <= tol |b| (exact where b is zero)."""
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b) <= tol * np.abs(b)


def subset(res, idx):
    """Contacts of pairs idx (in order) from a CSR result."""
    ints, reals, counts = [], [], []
    for p in idx:
        a, b = res.offsets[p], res.offsets[p + 1]
        ints.append(res.ints[a:b])
        reals.append(res.reals[a:b])
        counts.append(b - a)
    return np.concatenate(ints), np.concatenate(reals), np.array(counts)


def setup(ctx, s):
    ctx.upload_trees(s.trees)
    ctx.upload_bodies(s.tree_of_body, s.records)
    cl, k = scenes.clusters(s.pairs, s.n_bodies, s.records)
    ctx.upload_pairs(s.pairs)
    ctx.upload_clusters(cl, k)
    return cl, k


def test_batched_equals_per_scene_and_oracle():
    s = small_batch()
    print(f"[batched] bodies={s.n_bodies} pairs={s.pairs.shape[0]} scenes={len(s.counts)}", flush=True)
    ctx = d2.Context(0, d2.D2_FP64)
    try:
        cl, k = setup(ctx, s)
        n, steps = ctx.step_run(scenes.DT, True, want_steps=True)
        batched = ctx.result(with_pairs=True)
        bodies = ctx.download_bodies()
        assert n > 0
        # against the oracle: cluster steps, then contacts on the states the device used
        osc = oracle_scene(s, s.records)
        want = np.array([osc.cluster_step(s.pairs[cl == j], np.float32(scenes.DT)) for j in range(k)])
        assert np.allclose(steps, want, rtol=1e-12, atol=0)
        o_off, o_ints, o_reals = oracle_scene(s, bodies).narrow(s.pairs, filter=True)
        assert np.array_equal(batched.offsets, o_off)
        assert np.array_equal(batched.ints, o_ints)
        assert close(batched.reals, o_reals).all()
        # every scene alone gives the same contacts and steps
        sob = s.scene_of_body[s.pairs[:, 0]]
        for sc in range(len(s.counts)):
            idx = np.nonzero(sob == sc)[0]
            if idx.size == 0:
                continue
            used, lcl = np.unique(cl[idx], return_inverse=True)
            ctx.upload_bodies(s.tree_of_body, s.records)
            ctx.upload_pairs(s.pairs[idx])
            ctx.upload_clusters(lcl.astype(np.int32), used.size)
            _, st = ctx.step_run(scenes.DT, True, want_steps=True)
            alone = ctx.result(with_pairs=True)
            bi, br, bc = subset(batched, idx)
            assert np.array_equal(np.diff(alone.offsets), bc), f"scene {sc}: counts"
            ai = alone.ints.copy()
            ai[:, 0] = idx[ai[:, 0]]  # pair index is local to the upload
            assert np.array_equal(ai, bi), f"scene {sc}: index tuples"
            assert np.array_equal(alone.reals, br), f"scene {sc}: reals"
            assert np.array_equal(st, steps[used]), f"scene {sc}: cluster steps"
            assert np.array_equal(alone.per_pair, batched.per_pair[idx])
    finally:
        ctx.close()


def test_batched_fp32_graph_equals_fp64_eager():
    s = small_batch()
    out = {}
    for prec, graph in ((d2.D2_FP64, False), (d2.D2_FP32, True)):
        print(f"[batched] precision={prec} graph={graph}", flush=True)
        ctx = d2.Context(0, prec)
        try:
            setup(ctx, s)
            ctx.set_graph(graph)
            res = []
            for _ in range(3):  # observe, capture, replay
                ctx.upload_bodies(s.tree_of_body, s.records)
                _, st = ctx.step_run(scenes.DT, True, want_steps=True)
                res.append((st, ctx.result(with_pairs=True)))
            out[prec] = res
        finally:
            ctx.close()
    ref_st, ref = out[d2.D2_FP64][0]
    for st, r in out[d2.D2_FP32] + out[d2.D2_FP64]:
        assert np.array_equal(st, ref_st)
        assert np.array_equal(r.offsets, ref.offsets) and np.array_equal(r.ints, ref.ints)
        assert np.array_equal(r.reals, ref.reals) and np.array_equal(r.per_pair, ref.per_pair)


def test_autotune_shape_and_result():
    s = small_batch()
    ctx = d2.Context(0, d2.D2_FP32)
    try:
        setup(ctx, s)
        ctx.set_graph(True)
        _, st0 = ctx.step_run(scenes.DT, True, want_steps=True)
        r0 = ctx.result()
        shape, ms = ctx.autotune("step", reps=1)
        print(f"[autotune] shape={shape} ms={ms:.3f}", flush=True)
        assert ms > 0.0
        assert ctx.get_tune() == shape
        ctx.upload_bodies(s.tree_of_body, s.records)
        _, st1 = ctx.step_run(scenes.DT, True, want_steps=True)
        r1 = ctx.result()
        assert np.array_equal(st0, st1)
        assert np.array_equal(r0.ints, r1.ints) and np.array_equal(r0.reals, r1.reals)
    finally:
        ctx.close()


@functools.lru_cache(maxsize=None)
def large_batch():
    return scenes.scene_c3(seed=1107, n_scenes=12, lo=256, hi=512, threads=4)


def test_large_pair_list_equals_scenes_alone():
    """More pairs than k_scatter scans in-kernel (4096): the batched step (scan
    launch before the scatter; the filtered offsets scanned at result time)
    equals its scenes run alone (each below 4096 pairs: the in-kernel scan),
    in FP32 with graph replay and in FP64."""
    s = large_batch()
    print(f"[batched-large] bodies={s.n_bodies} pairs={s.pairs.shape[0]} scenes={len(s.counts)}", flush=True)
    assert s.pairs.shape[0] > 4096
    sob = s.scene_of_body[s.pairs[:, 0]]
    for prec, graph in ((d2.D2_FP32, True), (d2.D2_FP64, False)):
        print(f"[batched-large] precision={prec} graph={graph}", flush=True)
        ctx = d2.Context(0, prec)
        try:
            cl, k = setup(ctx, s)
            ctx.set_graph(graph)
            for _ in range(2):  # FP32: capture, then replay
                ctx.upload_bodies(s.tree_of_body, s.records)
                n, steps = ctx.step_run(scenes.DT, True, want_steps=True)
            batched = ctx.result(with_pairs=True)
            assert n == batched.offsets[-1] and n > 0
            ctx.set_graph(False)
            for sc in range(2):  # two scenes alone (each below 4096 pairs)
                idx = np.nonzero(sob == sc)[0]
                assert 0 < idx.size <= 4096
                used, lcl = np.unique(cl[idx], return_inverse=True)
                ctx.upload_bodies(s.tree_of_body, s.records)
                ctx.upload_pairs(s.pairs[idx])
                ctx.upload_clusters(lcl.astype(np.int32), used.size)
                _, st = ctx.step_run(scenes.DT, True, want_steps=True)
                alone = ctx.result(with_pairs=True)
                bi, br, bc = subset(batched, idx)
                assert np.array_equal(np.diff(alone.offsets), bc), f"scene {sc}: counts"
                ai = alone.ints.copy()
                ai[:, 0] = idx[ai[:, 0]]
                assert np.array_equal(ai, bi), f"scene {sc}: index tuples"
                assert np.array_equal(alone.reals, br), f"scene {sc}: reals"
                assert np.array_equal(st, steps[used]), f"scene {sc}: cluster steps"
        finally:
            ctx.close()
