"""GPU: the allocation-free bucket entry (d2_find_contacts_bucket_noalloc)
returns exactly what d2_find_contacts_bucket returns, on two calls in a row
(the second with a smaller input, so stale scratch contents would show), and
refuses inputs beyond its fixed scratch size before touching the device.
Named to be collected last: it must pass alone and after every other test."""
import numpy as np
import pytest

import delta2_amd as d2
from delta2_amd import scenes
from gpu_cases import dense_scene, with_future

pytestmark = pytest.mark.gpu
NOALLOC = "d2_find_contacts_bucket_noalloc"
MAX_NODE_PAIRS = 32768


def _node_pairs(s, a, b, limit, rng):
    """`limit` node pairs (node of a's tree, node of b's tree) drawn without
    repetition, in index order."""
    na = s.trees[s.tree_of_body[a]][0].shape[0]
    nb = s.trees[s.tree_of_body[b]][0].shape[0]
    grid = np.stack(np.meshgrid(np.arange(na), np.arange(nb), indexing="ij"), axis=-1).reshape(-1, 2)
    pick = rng.permutation(grid.shape[0])[:limit]
    return grid[np.sort(pick)].astype(np.int32)


def _same(got, want):
    for g, w in zip(got, want):  # pair_used, ints, reals: integer and FP64 results bit for bit
        assert g.shape == w.shape and np.array_equal(g, w)


def test_noalloc_bucket_equals_allocating_entry():
    s = dense_scene()
    rec = with_future(s.records, scenes.DT)
    rng = np.random.default_rng(3)
    with d2.Context(0) as c:
        c.upload_trees(s.trees)
        c.upload_bodies(s.tree_of_body, rec)
        # 1500 node pairs (at most 64 hits each: within the 131072-hit
        # scratch), then the first 300 of them: fewer hits, so leftovers of
        # the first call in the shared scratch would show
        a, b = (int(x) for x in s.pairs[0])
        first = _node_pairs(s, a, b, 1500, rng)
        hits = []
        for nps in (first, first[:300]):
            want = c.find_contacts_bucket(a, b, nps)
            got = c.find_contacts_bucket(a, b, nps, entry=NOALLOC)
            print(f"[noalloc] pair ({a}, {b}): {nps.shape[0]} node pairs, {want[1].shape[0]} hits", flush=True)
            _same(got, want)
            hits.append(want[1].shape[0])
        assert hits[0] > hits[1] > 0
        # beyond the fixed maximum: refused with a capacity error
        big = np.zeros((MAX_NODE_PAIRS + 1, 2), dtype=np.int32)
        with pytest.raises(d2.CapacityError):
            c.find_contacts_bucket(a, b, big, entry=NOALLOC)
        # the entry still works after the refusal
        _same(c.find_contacts_bucket(a, b, first[:300], entry=NOALLOC), c.find_contacts_bucket(a, b, first[:300]))
