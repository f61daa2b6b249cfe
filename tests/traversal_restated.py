"""A second, structurally different restatement of the reference's tree
traversals, used to pin the oracle's breadth-first ones (test infrastructure).

The oracle walks the trees breadth-first with explicit level lists, as
collision_detection/full_tree_comparison.h:13-219 does.  This module recurses
depth-first instead and restates the per-node-pair bucket functions in Python
on top of the oracle's primitives (comparisonDist, triTriCCDLinear,
pointTriCCDLinear, edgeEdgeCCD, transform -- each pinned by the reference's
own known-answer tests).  What it shares with the oracle is only those
primitives, so a mistake in the oracle's level bookkeeping (which side
advances, how children are routed to the soup or connected lists, the
emission order, the per-level min_time_out) shows up as a difference.

Emission order: the reference emits a level's contacts before the next
level's, and within a level in the order of the node-pair lists, which are
the children of the previous level's node pairs in parent order.  Every node
pair of a level lies the same number of expansions below the roots, so that
order is the lexicographic order of the child-index paths; within one level
the CCD's soup list (only the roots can be a leaf-leaf soup pair) precedes
the connected list.  Depth-first recursion visits paths in lexicographic
order, so sorting its contacts by (level, list, path) gives the reference's
order without a queue.
"""
import numpy as np

import d2o

class _Prims:
    """The oracle's FP64 CCD primitives called with arrays prepared once (the
    generic d2o wrappers convert every argument on every call)."""

    def __init__(self):
        self.L = d2o.lib()
        self.out = np.zeros(8)
        self.po = d2o._p(self.out)
        self.keep = []

    def arr(self, x):
        """A ctypes pointer to a float64 copy of x (kept alive with this object)."""
        a = np.ascontiguousarray(np.asarray(x, dtype=np.float64).ravel())
        self.keep.append(a)
        return d2o._p(a)

    def point_tri(self, pt, pTs, pTe, tri, tTs, tTe):
        self.L.d2o_point_tri_ccd_f64(pt, pTs, pTe, tri, tTs, tTe, self.po)
        o = self.out
        return o[0], o[1:4].copy(), o[4:7].copy(), o[7]

    def edge_edge(self, as_, ae, bs, be):
        self.L.d2o_edge_edge_ccd_f64(as_, ae, bs, be, self.po)
        o = self.out
        return o[0], o[1:4].copy(), o[4:7].copy(), o[7]

    def tri_tri(self, a, aTs, aTe, b, bTs, bTe):
        self.L.d2o_tri_tri_ccd_f64(a, aTs, aTe, b, bTs, bTe, self.po)
        o = self.out
        return o[0], o[1:4].copy(), o[4:7].copy(), o[7]


# node record fields (include/d2hip.h)
N_ID, N_BUCKET, N_INNER, N_DEPTH, N_PARENT, N_FIRST, N_NCHILD, N_NTRIS = range(8)


def bucket_tri(bucket, k):
    """Triangle k of an SoA bucket record as 9 doubles (A, B, C)."""
    b = bucket
    return np.array([b[k], b[8 + k], b[16 + k], b[24 + k], b[32 + k], b[40 + k], b[48 + k], b[56 + k], b[64 + k]])


def xform_tri(tri9, M):
    """Triangle::transformed (common/triangle.h:472-487) through the oracle's transform."""
    t = np.asarray(tri9).reshape(3, 3)
    return np.concatenate([d2o.transform_point(v, M) for v in t])


def tri_bbox(tri9):
    t = np.asarray(tri9).reshape(3, 3)
    return t.min(axis=0), t.max(axis=0)


def boxes_overlap(a, b):
    """bbox::overlap with strict intervals (common/bbox.h:8-46)."""
    (alo, ahi), (blo, bhi) = a, b
    return bool(np.all(alo < bhi) and np.all(blo < ahi))


class Body:
    def __init__(self, tree, record):
        self.nodes, self.buckets = tree
        self.rec = np.asarray(record, dtype=float)
        self.geo_eps = float(self.rec[37])

    def node(self, i):
        return self.nodes[i]

    def children(self, i):
        n = self.nodes[i]
        return [int(n[N_FIRST]) + c for c in range(int(n[N_NCHILD]))]

    def inner(self, i):
        return bool(self.nodes[i][N_INNER])

    def depth(self, i):
        return int(self.nodes[i][N_DEPTH])

    def tris(self, i):
        n = self.nodes[i]
        bk = self.buckets[int(n[N_BUCKET])]
        return [bucket_tri(bk, k) for k in range(int(n[N_NTRIS]))], bk[72:80], bk[80:88]


def advance_side(a, b, na, nb):
    """Which side's children a hitting node pair expands into
    (full_tree_comparison.h:31-49): the inner one; if both are inner, a when
    it is shallower, else b."""
    if not a.inner(na):
        return "b"
    if not b.inner(nb):
        return "a"
    return "a" if a.depth(na) < b.depth(nb) else "b"


# --- discrete: compareTreesFull -------------------------------------------
def bucket_discrete(a, b, na, nb, Ta, Tb):
    """findContactsBucketComparison for one node pair (comparison_dist.h:205-240):
    list of (kind 0, i, j, reals[11])."""
    search = float(np.float32(a.geo_eps + b.geo_eps))
    ta, eps_a, in_a = a.tris(na)
    tb, eps_b, in_b = b.tris(nb)
    wa = [xform_tri(t, Ta) for t in ta]
    wb = [xform_tri(t, Tb) for t in tb]
    out = []
    for i, A in enumerate(wa):
        for j, B in enumerate(wb):
            d, P, Q = d2o.comparison_dist(A, B)
            if d < (search + eps_a[i]) + eps_b[j]:
                r = np.concatenate([P, Q, [a.geo_eps + eps_a[i], b.geo_eps + eps_b[j], in_a[i], in_b[j], 0.0]])
                out.append((0, i, j, r))
    return out


def compare_trees_full(a, b, tested=None):
    """Depth-first restatement of compareTreesFull; returns (ints[n,8], reals[n,11])
    in the reference's order (pair column 0).  tested, a list, receives
    (level, node a, node b) of every node pair whose buckets are compared."""
    Ta = d2o.state_transformation(a.rec[14:28])
    Tb = d2o.state_transformation(b.rec[14:28])
    found = []

    def visit(na, nb, path):
        if tested is not None:
            tested.append((len(path), na, nb))
        hits = bucket_discrete(a, b, na, nb, Ta, Tb)
        if not hits:
            return
        if not a.inner(na) and not b.inner(nb):
            for kind, i, j, r in hits:
                found.append(((len(path), 0, path), (na, nb, kind, i, j, len(path)), r))
            return
        side = advance_side(a, b, na, nb)
        kids = a.children(na) if side == "a" else b.children(nb)
        for c, k in enumerate(kids):
            visit(k, nb, path + (c,)) if side == "a" else visit(na, k, path + (c,))

    visit(0, 0, ())
    return _pack(found)


# --- CCD: compareTreesFullContinuous --------------------------------------
def bucket_soup(a, b, na, nb, T):
    """findContactsBucketContinuousComparison for one node pair
    (continuous_comparison.h:783-828): swept boxes, then triTriCCDLinear;
    surrogate eps is not part of the threshold."""
    aTs, aTe, bTs, bTe = T
    search = float(np.float32(a.geo_eps + b.geo_eps))
    ta, _, _ = a.tris(na)
    tb, _, _ = b.tris(nb)

    def swept(t, Ts, Te, g):
        lo0, hi0 = tri_bbox(xform_tri(t, Ts))
        lo1, hi1 = tri_bbox(xform_tri(t, Te))
        return np.minimum(lo0, lo1) - g, np.maximum(hi0, hi1) + g

    F = _Prims()
    M = [F.arr(m) for m in T]
    bboxes = [swept(bt, bTs, bTe, b.geo_eps) for bt in tb]
    bptr = [F.arr(bt) for bt in tb]
    out = []
    for i, at in enumerate(ta):
        abox = swept(at, aTs, aTe, a.geo_eps)
        aptr = F.arr(at)
        for j in range(len(tb)):
            if not boxes_overlap(abox, bboxes[j]):
                continue
            d, P, Q, toc = F.tri_tri(aptr, M[0], M[1], bptr[j], M[2], M[3])
            if d < search:
                out.append((0, i, j, np.concatenate([P, Q, [a.geo_eps, b.geo_eps, 0.0, 0.0, toc]])))
    return out


def bucket_connected(a, b, na, nb, T):
    """findContactsBucketConnectedContinuousComparison for one leaf pair
    (continuous_comparison.h:830-910): every bucket vertex of a against every
    triangle of b, then of b against a (stored as (Q, P)), then every edge
    pair by edgeEdgeCCD (stored as (Q, P))."""
    F = _Prims()
    aTs, aTe, bTs, bTe = T
    M = [F.arr(m) for m in T]
    search = float(np.float32(a.geo_eps + b.geo_eps))
    ta, _, _ = a.tris(na)
    tb, _, _ = b.tris(nb)
    pa = [F.arr(t) for t in ta]
    pb = [F.arr(t) for t in tb]
    out = []

    def push(kind, i, j, A, B, toc):
        out.append((kind, i, j, np.concatenate([A, B, [a.geo_eps, b.geo_eps, 0.0, 0.0, toc]])))

    for v in range(3 * len(ta)):
        p = F.arr(ta[v // 3][3 * (v % 3):3 * (v % 3) + 3])
        for j, t in enumerate(pb):
            d, P, Q, toc = F.point_tri(p, M[0], M[1], t, M[2], M[3])
            if d < search:
                push(1, v, j, P, Q, toc)
    for v in range(3 * len(tb)):
        p = F.arr(tb[v // 3][3 * (v % 3):3 * (v % 3) + 3])
        for i, t in enumerate(pa):
            d, P, Q, toc = F.point_tri(p, M[2], M[3], t, M[0], M[1])
            if d < search:
                push(2, i, v, Q, P, toc)

    def edges(tris, Ts, Te):
        """Every bucket edge k of triangle t (t[k], t[k+1]) at both poses."""
        out = []
        for t in tris:
            v = np.asarray(t).reshape(3, 3)
            ws = [d2o.transform_point(x, Ts) for x in v]
            we = [d2o.transform_point(x, Te) for x in v]
            for k in range(3):
                out.append((F.arr(np.concatenate([ws[k], ws[(k + 1) % 3]])),
                            F.arr(np.concatenate([we[k], we[(k + 1) % 3]]))))
        return out

    eb_list = edges(tb, bTs, bTe)
    for ea, (as_, ae) in enumerate(edges(ta, aTs, aTe)):
        for eb, (bs, be) in enumerate(eb_list):
            d, P, Q, toc = F.edge_edge(as_, ae, bs, be)
            if d < search:
                push(3, ea, eb, Q, P, toc)
    return out


def ccd_frames(a, b, max_time):
    """Start and end poses: the current states and current.extrapolate(max_time)
    (Particle::projectFutureState, full_tree_comparison.h:78-79)."""
    fa = d2o.extrapolate(a.rec[0:14], a.rec[28:37], max_time)
    fb = d2o.extrapolate(b.rec[0:14], b.rec[28:37], max_time)
    return (d2o.state_transformation(a.rec[0:14]), d2o.state_transformation(fa),
            d2o.state_transformation(b.rec[0:14]), d2o.state_transformation(fb))


def compare_trees_full_continuous(a, b, max_time, tested=None):
    """Depth-first restatement of compareTreesFullContinuous; returns
    (ints, reals, min_time_out).  The roots are a soup pair whatever they are;
    below them a pair of two leaves is a connected pair, every other a soup
    pair.  min_time_out is the smallest toc x max_time over every hit of a
    level's pass in which some leaf-leaf node pair hit (full_tree_comparison.h:
    104-109, 165-170)."""
    T = ccd_frames(a, b, max_time)
    found = []
    pass_hits = {}  # (level, list) -> every hit's toc in that pass
    pass_leaf = set()  # passes in which a leaf-leaf node pair hit

    def visit(na, nb, path, root):
        if tested is not None:
            tested.append((len(path), na, nb))
        lst = 0 if (root or a.inner(na) or b.inner(nb)) else 1
        key = (len(path), lst)
        hits = (bucket_soup if lst == 0 else bucket_connected)(a, b, na, nb, T)
        pass_hits.setdefault(key, []).extend(r[10] for *_, r in hits)
        if not hits:
            return
        if not a.inner(na) and not b.inner(nb):
            pass_leaf.add(key)
            for kind, i, j, r in hits:
                found.append(((len(path), lst, path), (na, nb, kind, i, j, 2 * len(path) + lst), r))
            return
        side = advance_side(a, b, na, nb)
        kids = a.children(na) if side == "a" else b.children(nb)
        for c, k in enumerate(kids):
            visit(k, nb, path + (c,), False) if side == "a" else visit(na, k, path + (c,), False)

    visit(0, 0, (), True)
    min_time = max_time
    for key in pass_leaf:
        min_time = min([min_time] + [t * max_time for t in pass_hits[key]])
    ints, reals = _pack(found)
    return ints, reals, min_time


def _pack(found):
    found.sort(key=lambda f: f[0])
    ints = np.zeros((len(found), 8), dtype=np.int32)
    reals = np.zeros((len(found), 11))
    for k, (_, (na, nb, kind, i, j, level), r) in enumerate(found):
        ints[k] = (0, na, nb, kind, i, j, level, 0)
        reals[k] = r
    return ints, reals


# --- exhaustive leaf pairs -------------------------------------------------
def leaves(body):
    return [i for i in range(body.nodes.shape[0]) if not body.inner(i)]


def exhaustive_discrete(a, b):
    """Every leaf of a against every leaf of b: {(na, nb, kind, i, j): reals}."""
    Ta = d2o.state_transformation(a.rec[14:28])
    Tb = d2o.state_transformation(b.rec[14:28])
    out = {}
    for la in leaves(a):
        for lb in leaves(b):
            for kind, i, j, r in bucket_discrete(a, b, la, lb, Ta, Tb):
                out[(la, lb, kind, i, j)] = r
    return out


def exhaustive_connected(a, b, max_time):
    T = ccd_frames(a, b, max_time)
    out = {}
    for la in leaves(a):
        for lb in leaves(b):
            for kind, i, j, r in bucket_connected(a, b, la, lb, T):
                out[(la, lb, kind, i, j)] = r
    return out
