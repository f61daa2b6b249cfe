"""GPU: the algorithmic-byte accounting (d2_set_touch_stats) counts exactly
the (body, node) records the reference traversal tests.

The expected sets come from the depth-first restatement of the traversals
(tests/traversal_restated.py), which records every node pair whose buckets
it compares; per level and over the run, the distinct inner nodes, leaf
nodes, their triangles and the bodies must equal the device's counts, in the
FP64 path and in the FP32 form (which culls triangle pairs but tests the same
node pairs).
"""
import numpy as np
import pytest

import delta2_amd as d2
import traversal_restated as tr
from delta2_amd import scenes
from gpu_cases import dense_scene, hull_scene, with_future

pytestmark = pytest.mark.gpu


def expected(s, pairs, records, continuous, max_time):
    levels, union_nodes, union_bodies = {}, set(), set()
    for a, b in pairs:
        A = tr.Body(s.trees[s.tree_of_body[a]], records[a])
        B = tr.Body(s.trees[s.tree_of_body[b]], records[b])
        tested = []
        if continuous:
            tr.compare_trees_full_continuous(A, B, max_time, tested)
        else:
            tr.compare_trees_full(A, B, tested)
        for lvl, na, nb in tested:
            nodes, bodies = levels.setdefault(lvl, (set(), set()))
            nodes.update([(int(a), int(na)), (int(b), int(nb))])
            bodies.update([int(a), int(b)])
            union_nodes.update([(int(a), int(na)), (int(b), int(nb))])
            union_bodies.update([int(a), int(b)])

    def count(nodes, bodies):
        inner = leaf = tris = 0
        for body, n in nodes:
            rec = s.trees[s.tree_of_body[body]][0][n]
            if rec[2]:
                inner += 1
            else:
                leaf += 1
                tris += int(rec[7])
        return [inner, leaf, tris, len(bodies)]

    per_level = np.zeros((17, 4), dtype=np.int64)
    for lvl, (nodes, bodies) in levels.items():
        per_level[lvl] = count(nodes, bodies)
    return per_level, np.array(count(union_nodes, union_bodies))


@pytest.mark.parametrize("make", [dense_scene, hull_scene], ids=["dense", "hulls"])
@pytest.mark.parametrize("prec", [d2.D2_FP64, d2.D2_FP32], ids=["fp64", "fp32"])
def test_touched_nodes_match_the_reference_traversal(make, prec):
    s = make()
    pairs = s.pairs[:24]
    mt = 2 * scenes.DT
    with d2.Context(0, prec) as c:
        c.upload_trees(s.trees)
        c.set_touch_stats(True)
        rec = with_future(s.records, mt)
        c.upload_bodies(s.tree_of_body, rec)
        c.upload_pairs(pairs)
        c.narrow_run(False)
        got_lv, got_run = c.touch_stats()
        want_lv, want_run = expected(s, pairs, rec, False, mt)
        print(f"[touch] {s.name} precision={prec} discrete run {got_run.tolist()}", flush=True)
        assert np.array_equal(got_lv[0], want_lv), (got_lv[0], want_lv)
        assert np.array_equal(got_run, want_run)
        c.upload_bodies(s.tree_of_body, s.records)
        c.upload_pairs(pairs, np.full(pairs.shape[0], mt))
        c.ccd_run()
        got_lv, got_run = c.touch_stats()
        want_lv, want_run = expected(s, pairs, s.records, True, mt)
        print(f"[touch] {s.name} precision={prec} ccd run {got_run.tolist()}", flush=True)
        assert np.array_equal(got_lv[1], want_lv), (got_lv[1], want_lv)
        assert np.array_equal(got_run, want_run)
        assert want_run[0] > 0 and want_run[1] > 0  # inner and leaf nodes were tested
