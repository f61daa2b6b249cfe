"""GPU: call orders of the C ABI and what each upload keeps or forgets.

- A tree upload forgets bodies, pairs, clusters and groups (body records name
  trees by index): runs fail with D2_ERR_STATE until they come again.
- A body upload with another count forgets pairs and clusters; with the same
  count it keeps them and the scene groups.
- The learned gate's certificate stays sound for trees whose leaf buckets
  carry eps (the discrete hit test adds them, comparison_dist.h:233).
"""
import numpy as np
import pytest

import delta2_amd as d2
import mlp_oracle as mo
from delta2_amd import scenes
from gpu_cases import dense_scene, with_future

pytestmark = pytest.mark.gpu


def _state_error(fn):
    with pytest.raises(d2.D2Error) as e:
        fn()
    assert e.value.code == d2.D2_ERR_STATE, e.value


def test_tree_upload_forgets_bodies_pairs_and_clusters():
    s = dense_scene()
    with d2.Context(0) as c:
        c.upload_trees(s.trees)
        c.upload_bodies(s.tree_of_body, s.records)
        c.upload_pairs(s.pairs, np.full(s.pairs.shape[0], scenes.DT))
        c.upload_clusters(np.zeros(s.pairs.shape[0], np.int32), 1)
        assert c.ccd_run() > 0
        c.upload_trees(s.trees[:3])  # fewer trees: the kept bodies would name missing ones
        _state_error(lambda: c.ccd_run())
        _state_error(lambda: c.narrow_run(True))
        _state_error(lambda: c.step_run(scenes.DT))
        _state_error(lambda: c.upload_pairs(s.pairs))  # no bodies
        _state_error(lambda: c.broad_phase(1))
        c.upload_trees(s.trees)
        c.upload_bodies(s.tree_of_body, s.records)
        _state_error(lambda: c.ccd_run())  # no pairs yet
        _state_error(lambda: c.upload_clusters(np.zeros(1, np.int32), 1))
        c.upload_pairs(s.pairs, np.full(s.pairs.shape[0], scenes.DT))
        assert c.ccd_run() > 0


def test_body_upload_of_another_count_forgets_pairs():
    s = dense_scene()
    rec = with_future(s.records, scenes.DT)
    with d2.Context(0) as c:
        c.upload_trees(s.trees)
        c.upload_bodies(s.tree_of_body, rec)
        c.upload_pairs(s.pairs)
        want = c.narrow_run(True)
        # the same bodies again: pairs stay
        c.upload_bodies(s.tree_of_body, rec)
        assert c.narrow_run(True) == want
        # fewer bodies: the pair list (which names bodies past the end) is gone
        c.upload_bodies(s.tree_of_body[:10], rec[:10])
        _state_error(lambda: c.narrow_run(True))
        _state_error(lambda: c.build_clusters())


def test_body_upload_of_the_same_count_keeps_groups():
    """Groups split one body list into scenes; uploading the same bodies again
    (as a stepping loop does) keeps them, so no pair crosses scenes."""
    s = dense_scene()
    n = s.n_bodies
    groups = (np.arange(n) >= n // 2).astype(np.int32)
    with d2.Context(0) as c:
        c.upload_trees(s.trees)
        c.upload_bodies(s.tree_of_body, s.records)
        c.upload_groups(groups)
        n_grouped = c.broad_phase(1)
        pairs, _ = c.download_pairs()
        assert np.all(groups[pairs[:, 0]] == groups[pairs[:, 1]])
        c.upload_bodies(s.tree_of_body, s.records)
        assert c.broad_phase(1) == n_grouped
        pairs, _ = c.download_pairs()
        assert np.all(groups[pairs[:, 0]] == groups[pairs[:, 1]])
        c.upload_groups(None)
        assert c.broad_phase(1) > n_grouped  # the scene as one group has cross pairs


def test_certificate_counts_leaf_bucket_eps():
    """Leaf buckets with eps > 0 widen the discrete hit test by eps_a + eps_b;
    a network that proposes every pair leaves the certificate to decide, and
    the results must equal the plain path's (the eps makes far pairs hit)."""
    s = dense_scene()
    trees = []
    for nodes, buckets in s.trees:
        b = buckets.copy()
        leaf = nodes[:, 2] == 0
        b[leaf, 72:80] = np.where(np.arange(8)[None, :] < nodes[leaf, 7:8], 0.3, 0.0)
        trees.append((nodes, b))
    rec = with_future(s.records, scenes.DT)
    p = np.zeros(mo.N_PARAMS, dtype=np.float32)
    p[-2:] = (1.0, 2.0)  # every pair proposed as contact-free
    # a longer pair list: the broad phase with inflated radii reaches farther pairs
    pairs = scenes.broad_phase(rec, s.radii + 0.6)
    for prec in (d2.D2_FP64, d2.D2_FP32):
        with d2.Context(0, prec) as c:
            c.upload_trees(trees)
            c.upload_bodies(s.tree_of_body, rec)
            plain = d2.compare_trees_full(c, pairs, filter=False)
            c.mlp_set_params(p)
            c.set_learned(True, phases=2)
            gated = d2.compare_trees_full(c, pairs, filter=False)
            proposed, certified = c.gate_counts()["discrete"]
            print(f"[cert leaf eps] precision={prec} pairs={pairs.shape[0]} contacts={plain.n} "
                  f"proposed={proposed} certified={certified}", flush=True)
            assert np.array_equal(plain.offsets, gated.offsets)
            assert np.array_equal(plain.ints, gated.ints) and np.array_equal(plain.reals, gated.reals)
            # pairs whose contacts come only from the leaf eps exist
            assert plain.n > 0 and proposed == pairs.shape[0]
