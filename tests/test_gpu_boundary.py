"""GPU: the plugin point's bucket comparison and the standalone contact filter
through the C ABI.

- d2_find_contacts_bucket (ContactDetectionStrategy::findContactsBucketComparison,
  comparison_dist.h:205-240) equals the Python restatement of the bucket
  comparison on the oracle's primitives, bit for bit, with the reference's
  cumulative pair_used_out; the reference fixture (test_comparison.cpp:90-128:
  8 hits); capacity and argument errors.
- d2_filter_contacts (filterContacts, contact.h:44-69) returns the oracle's
  std::sort permutation and kept set, on tie-heavy inputs of every size the
  finishing kernels route differently (up to 512, up to 2048, above).
"""
import numpy as np
import pytest

import d2o
import delta2_amd as d2
import traversal_restated as tr
from delta2_amd import scenes
from gpu_cases import dense_scene, hull_scene, with_future

pytestmark = pytest.mark.gpu


def _node_pairs(s, a, b, rng, n):
    """A mixed list: the roots, every (root child, root child) pair, random
    node pairs of both trees (inner and leaf, repeated entries included)."""
    na = s.trees[s.tree_of_body[a]][0].shape[0]
    nb = s.trees[s.tree_of_body[b]][0].shape[0]
    pairs = [(0, 0)] + [(i, j) for i in range(1, min(9, na)) for j in range(1, min(9, nb))]
    pairs += list(zip(rng.integers(0, na, n), rng.integers(0, nb, n)))
    pairs.append(pairs[3])
    return np.array(pairs, dtype=np.int32)


@pytest.mark.parametrize("make", [dense_scene, hull_scene], ids=["dense", "hulls"])
def test_bucket_entry_matches_restatement(make):
    s = make()
    rec = with_future(s.records, scenes.DT)
    rng = np.random.default_rng(7)
    with d2.Context(0) as c:
        c.upload_trees(s.trees)
        c.upload_bodies(s.tree_of_body, rec)
        checked = 0
        for p in range(0, min(12, s.pairs.shape[0])):
            a, b = (int(x) for x in s.pairs[p])
            nps = _node_pairs(s, a, b, rng, 300)
            used, ints, reals = c.find_contacts_bucket(a, b, nps)
            A = tr.Body(s.trees[s.tree_of_body[a]], rec[a])
            B = tr.Body(s.trees[s.tree_of_body[b]], rec[b])
            Ta = d2o.state_transformation(rec[a, 14:28])
            Tb = d2o.state_transformation(rec[b, 14:28])
            want_i, want_r, want_used = [], [], []
            for na, nb in nps:
                for kind, i, j, r in tr.bucket_discrete(A, B, int(na), int(nb), Ta, Tb):
                    want_i.append((0, na, nb, kind, i, j, 0, 0))
                    want_r.append(r)
                want_used.append(len(want_i))
            assert np.array_equal(used, want_used), p
            assert np.array_equal(ints, np.array(want_i, dtype=np.int32).reshape(-1, 8)), p
            assert np.array_equal(reals, np.array(want_r).reshape(-1, 11)), p
            checked += len(want_i)
        print(f"[bucket] {s.name}: {checked} hits checked", flush=True)
        assert checked > 100


def test_bucket_entry_reference_fixture_and_errors():
    """test_comparison.cpp:90-128: the pyramid's 4 triangles against the
    plane's 2 at 0.105 above: 8 hits in one node pair."""
    from test_oracle_kat import PLANE_F, PLANE_V, PYRAMID_F, PYRAMID_V, body_record
    rng = d2.pcg_seed(42, 54)
    trees = [d2.tree_build(PYRAMID_V, PYRAMID_F, rng), d2.tree_build(PLANE_V, PLANE_F, rng)]
    bodies = np.stack([body_record(fut_trans=(0, 0, 1.0), static=True), body_record(fut_trans=(0, 0, 1.105))])
    with d2.Context(0) as c:
        c.upload_trees(trees)
        c.upload_bodies([0, 1], bodies)
        used, ints, reals = c.find_contacts_bucket(0, 1, [[0, 0]])
        assert used.tolist() == [8] and ints.shape == (8, 8)
        used, ints, reals = c.find_contacts_bucket(0, 1, [[0, 0], [0, 0]])
        assert used.tolist() == [8, 16]
        used, ints, reals = c.find_contacts_bucket(0, 1, np.zeros((0, 2), np.int32))
        assert used.size == 0 and ints.shape == (0, 8)
        with pytest.raises(d2.CapacityError) as e:
            c.find_contacts_bucket(0, 1, [[0, 0]], max_hits=5)
        assert c.last_hits == 8
        with pytest.raises(d2.D2Error) as e:
            c.find_contacts_bucket(0, 1, [[0, 1]])  # the plane tree has one node
        assert e.value.code == d2.D2_ERR_ARG


def _tie_heavy(rng, n):
    """Contacts with few distinct eps sums and distances (exact ties) and
    midpoints packed closely enough that the greedy pass rejects many."""
    A = rng.integers(0, 6, (n, 3)) * 0.01
    d = rng.integers(0, 3, (n, 1)) * 0.005 * np.array([[0.0, 0.0, 1.0]])
    eps = rng.integers(1, 3, (n, 2)) * 0.01
    return np.concatenate([A, A + d, eps, np.zeros((n, 3))], axis=1)


@pytest.mark.parametrize("n", [1, 7, 300, 512, 513, 1500, 2048, 2600])
def test_filter_entry_matches_oracle(n):
    rng = np.random.default_rng(n)
    reals = _tie_heavy(rng, n)
    with d2.Context(0) as c:
        for margin in (0.0, 0.004):
            order, kept = c.filter_contacts(reals, margin)
            want_order, want_kept = d2o.filter_order(reals, margin)
            assert np.array_equal(order, want_order), (n, margin)
            assert np.array_equal(kept, want_kept), (n, margin)
            assert 0 < kept.size <= n


def test_filter_entry_reference_fixture():
    """test_comparison.cpp:90-128: the 8 pyramid-plane contacts filter to 1."""
    from test_oracle_kat import PLANE_F, PLANE_V, PYRAMID_F, PYRAMID_V, body_record
    rng = d2.pcg_seed(42, 54)
    trees = [d2.tree_build(PYRAMID_V, PYRAMID_F, rng), d2.tree_build(PLANE_V, PLANE_F, rng)]
    bodies = np.stack([body_record(fut_trans=(0, 0, 1.0), static=True), body_record(fut_trans=(0, 0, 1.105))])
    with d2.Context(0) as c:
        c.upload_trees(trees)
        c.upload_bodies([0, 1], bodies)
        _, _, reals = c.find_contacts_bucket(0, 1, [[0, 0]])
        order, kept = c.filter_contacts(reals, 0.0)
        assert sorted(order.tolist()) == list(range(8)) and kept.size == 1
        empty_order, empty_kept = c.filter_contacts(np.zeros((0, 11)), 0.0)
        assert empty_order.size == 0 and empty_kept.size == 0
