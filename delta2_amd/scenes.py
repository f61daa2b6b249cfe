"""Synthetic scenes for the contact-detection path (inputs only).

Builds the meshes, body records, surrogate trees and broad-phase pair lists of
the workloads the path is measured on.  Trees come from the library's own
``tree_build`` (the reference builds them in ``MeshData``); everything else
is plain numpy/scipy input generation.

Conventions (stated once, used everywhere):
  * the tree-building PCG32 is seeded (42, 54) and shared by all meshes in
    body order, as one ``Options`` object is in the reference
    (common/cli_options.h:46, runner/main.cpp:38);
  * scene randomness uses numpy's PCG64 seeded 1000 + workload number;
  * geo_eps = 0.01 (common/cli_options.h:41), density 1, angular momentum 0,
    time step dt = float(1/60) (Options::time_step_size is a float);
  * velocities "N(0, s)" are normal with standard deviation s per axis;
  * trees and bounding radii are built from the un-centred vertices, the
    inertia about the centre of mass (model/mesh.cpp:20-30 builds the tree
    at :21; initInertiaDataVolume re-centres the vertices at :75).
"""
import math

import numpy as np

from . import pcg_seed, tree_build

DT = float(np.float32(1.0 / 60.0))
GEO_EPS = 0.01


# --- meshes -----------------------------------------------------------------
def icosphere(subdiv, radius=1.0):
    t = (1.0 + 5.0 ** 0.5) / 2.0
    V = [[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
         [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]]
    F = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
         [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11],
         [6, 2, 10], [8, 6, 7], [9, 8, 1]]
    V = [np.array(v, dtype=float) / np.linalg.norm(v) for v in V]
    for _ in range(subdiv):
        mid = {}

        def m(a, b):
            k = (min(a, b), max(a, b))
            if k not in mid:
                p = V[a] + V[b]
                V.append(p / np.linalg.norm(p))
                mid[k] = len(V) - 1
            return mid[k]

        F2 = []
        for a, b, c in F:
            ab, bc, ca = m(a, b), m(b, c), m(c, a)
            F2 += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        F = F2
    return np.array(V) * radius, np.array(F, dtype=np.int32)


def convex_hull_mesh(points):
    """Outward-oriented triangle mesh of the convex hull (Qhull's quickhull)."""
    from scipy.spatial import ConvexHull
    h = ConvexHull(points)
    used = np.unique(h.simplices)
    remap = -np.ones(points.shape[0], dtype=np.int64)
    remap[used] = np.arange(used.size)
    V = points[used]
    F = remap[h.simplices]
    c = V.mean(axis=0)
    n = np.cross(V[F[:, 1]] - V[F[:, 0]], V[F[:, 2]] - V[F[:, 0]])
    flip = np.einsum("ij,ij->i", n, V[F[:, 0]] - c) < 0
    F[flip] = F[flip][:, [0, 2, 1]]
    return V, F.astype(np.int32)


def ellipsoid_hull(rng, n_points=64, lo=0.3, hi=1.2):
    axes = rng.uniform(lo, hi, size=3)
    d = rng.normal(size=(n_points, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return convex_hull_mesh(d * axes)


def mass_properties(V, F):
    """Volume, centre of mass and unit-density inertia about it (divergence theorem
    over the tetrahedra (0, a, b, c)); replaces the reference's tetgen route
    (model/mesh.cpp:53-91), which is absent here."""
    a, b, c = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    det = np.einsum("ij,ij->i", a, np.cross(b, c))
    vol = det.sum() / 6.0
    com = (det[:, None] * (a + b + c)).sum(axis=0) / (24.0 * vol)
    canon = np.array([[2.0, 1, 1], [1, 2, 1], [1, 1, 2]]) / 120.0
    A = np.stack([a, b, c], axis=2)  # columns a, b, c
    C = np.einsum("n,nij,jk,nlk->il", det, A, canon, A)
    C = C - vol * np.outer(com, com)
    inertia = np.trace(C) * np.eye(3) - C
    return vol, com, inertia


def random_rotation(rng):
    """Uniform unit quaternion (w, x, y, z), Shoemake's method."""
    u1, u2, u3 = rng.uniform(size=3)
    q = np.array([math.sqrt(u1) * math.cos(2 * math.pi * u3), math.sqrt(1 - u1) * math.sin(2 * math.pi * u2),
                  math.sqrt(1 - u1) * math.cos(2 * math.pi * u2), math.sqrt(u1) * math.sin(2 * math.pi * u3)])
    return q


# --- scenes -----------------------------------------------------------------
class Scene:
    """Bodies (records, 40 doubles each), their trees, and a broad-phase pair list."""

    def __init__(self, name, meshes, trees, records, radii, scene_of_body=None):
        self.name = name
        self.meshes = meshes
        self.trees = trees
        self.records = records
        self.radii = radii
        self.tree_of_body = np.arange(len(trees), dtype=np.int32)
        self.scene_of_body = np.zeros(len(trees), dtype=np.int32) if scene_of_body is None else scene_of_body
        self.pairs = np.zeros((0, 2), dtype=np.int32)
        self._n_tris = None

    @property
    def n_bodies(self):
        return self.records.shape[0]

    def n_triangles(self):
        if self._n_tris is None:
            self._n_tris = int(sum(F.shape[0] for _, F in self.meshes))
        return self._n_tris

    def save(self, path):
        t = self.trees if isinstance(self.trees, FlatTrees) else FlatTrees.from_list(list(self.trees))
        np.savez(path, name=self.name, offs=t.offs, nodes=t.nodes, buckets=t.buckets, records=self.records,
                 radii=self.radii, scene_of_body=self.scene_of_body, pairs=self.pairs, n_tris=self.n_triangles(),
                 box=getattr(self, "box", 0.0))

    @staticmethod
    def load(path):
        z = np.load(path)
        trees = FlatTrees(z["offs"].astype(np.int32), z["nodes"], z["buckets"])
        s = Scene(str(z["name"]), [], trees, z["records"], z["radii"], z["scene_of_body"])
        s.pairs = z["pairs"]
        s._n_tris = int(z["n_tris"])
        s.box = float(z["box"])
        return s


def cached(name, build):
    """build() unless $D2_SCENE_CACHE holds <name>.npz (then load it; save after building)."""
    import os
    d = os.environ.get("D2_SCENE_CACHE")
    if not d:
        return build()
    path = os.path.join(d, name + ".npz")
    if os.path.exists(path):
        return Scene.load(path)
    s = build()
    os.makedirs(d, exist_ok=True)
    s.save(path)
    return s


def body_record(q, pos, vel, inv_inertia, geo_eps=GEO_EPS, static=False, last_step=DT):
    r = np.zeros(40)
    r[0:4] = q
    r[4:7] = pos
    r[7:10] = vel
    r[14:28] = r[0:14]
    r[28:37] = np.asarray(inv_inertia).ravel()
    r[37] = geo_eps
    r[38] = 1.0 if static else 0.0
    r[39] = last_step
    return r


def _bodies(meshes, rng_scene, positions, vel_sigma, tree_rng):
    trees, records, radii = [], [], []
    for k, (V, F) in enumerate(meshes):
        trees.append(tree_build(V, F, tree_rng))
        vol, com, inertia = mass_properties(V, F)
        inv = np.linalg.inv(inertia)  # density 1
        q = random_rotation(rng_scene)
        v = rng_scene.normal(0.0, vel_sigma, size=3)
        records.append(body_record(q, positions[k], v, inv))
        radii.append(float(np.max(np.linalg.norm(V, axis=1))))
    return trees, np.array(records), np.array(radii)


def poisson_disk(rng, radii, box, max_tries=2000):
    """Random sequential placement of non-overlapping spheres (largest first)."""
    n = len(radii)
    order = np.argsort(-np.asarray(radii), kind="stable")
    rmax = float(np.max(radii))
    cell = 2.0 * rmax
    nc = max(1, int(box / cell))
    cell = box / nc
    grid = {}
    pos = np.zeros((n, 3))
    for i in order:
        r = radii[i]
        for _ in range(max_tries):
            p = rng.uniform(r, box - r, size=3)
            ci = tuple(np.minimum((p / cell).astype(int), nc - 1))
            ok = True
            for dx in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for dz in (-1, 0, 1):
                        for j in grid.get((ci[0] + dx, ci[1] + dy, ci[2] + dz), ()):
                            if np.sum((pos[j] - p) ** 2) < (r + radii[j]) ** 2:
                                ok = False
                                break
                        if not ok:
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if ok:
                pos[i] = p
                grid.setdefault(ci, []).append(i)
                break
        else:
            raise RuntimeError("poisson_disk: packing too dense for random placement")
    return pos


def clusters(pairs, n_bodies, records=None):
    """Connected components of the pair graph, as separate_clusters.cpp:355-459
    forms clusters (static bodies do not join clusters together).  Returns
    (cluster_of_pair, n_clusters) with clusters numbered by first pair."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components
    pairs = np.asarray(pairs, dtype=np.int64).reshape(-1, 2)
    if pairs.shape[0] == 0:
        return np.zeros(0, dtype=np.int32), 1
    a, b = pairs[:, 0].copy(), pairs[:, 1].copy()
    if records is not None:  # a static body links nothing: give each of its pairs a private node
        static = records[:, 38] != 0
        extra = n_bodies + np.arange(pairs.shape[0])
        a = np.where(static[a], extra, a)
        b = np.where(static[b], extra, b)
    n = n_bodies + pairs.shape[0]
    g = coo_matrix((np.ones(a.size), (a, b)), shape=(n, n))
    _, label = connected_components(g, directed=False)
    lab = label[a]  # a pair's endpoints share a component
    _, first = np.unique(lab, return_index=True)
    order = np.argsort(first)
    remap = np.empty(order.size, dtype=np.int32)
    remap[order] = np.arange(order.size, dtype=np.int32)
    _, inv = np.unique(lab, return_inverse=True)
    return remap[inv].astype(np.int32), int(order.size)


def broad_phase(records, radii, geo_eps=GEO_EPS, scene_of_body=None):
    """Pairs whose bounding spheres (radius + geo_eps) meet within the step, by
    collideFuncIndividualTime's rule (collision_detection/broad_phase_embree.cpp:
    208-229 with continuous_comparison.h:12-36): t >= 0 and t < max(last steps).
    Returns (a, b) with a < b, sorted; pairs never cross scenes."""
    from scipy.spatial import cKDTree
    pos = records[:, 4:7]
    vel = records[:, 7:10]
    last = records[:, 39]
    R = radii + geo_eps
    reach = 2.0 * R.max() + 2.0 * np.linalg.norm(vel, axis=1).max() * last.max() + 1e-9
    cand = cKDTree(pos).query_pairs(reach, output_type="ndarray")
    if cand.size == 0:
        return np.zeros((0, 2), dtype=np.int32)
    a, b = np.minimum(cand[:, 0], cand[:, 1]), np.maximum(cand[:, 0], cand[:, 1])
    if scene_of_body is not None:
        keep = scene_of_body[a] == scene_of_body[b]
        a, b = a[keep], b[keep]
    static = records[:, 38] != 0
    keep = ~(static[a] & static[b])
    a, b = a[keep], b[keep]
    ab = pos[a] - pos[b]
    rv = vel[a] - vel[b]
    r = R[a] + R[b]
    c = np.einsum("ij,ij->i", ab, ab) - r * r
    qa = np.einsum("ij,ij->i", rv, rv)
    qb = np.einsum("ij,ij->i", rv, ab)
    d = qb * qb - qa * c
    with np.errstate(invalid="ignore", divide="ignore"):
        t = np.where(c < 0.0, 0.0, np.where(qb >= 0.0, -1.0, np.where(d < 0.0, -1.0, (-qb - np.sqrt(np.maximum(d, 0))) / qa)))
    hit = (t >= 0.0) & (t < np.maximum(last[a], last[b]))
    out = np.stack([a[hit], b[hit]], axis=1).astype(np.int32)
    order = np.lexsort((out[:, 1], out[:, 0]))
    return out[order]


def scene_c1(seed=1001, n_bodies=100, subdiv=2, box=10.0):
    """100 subdivision-2 icospheres scaled U(0.5,1.5), random rotations,
    positions U([0,box]^3), velocities N(0,1)."""
    rng = np.random.default_rng(seed)
    meshes = []
    for _ in range(n_bodies):
        V, F = icosphere(subdiv, rng.uniform(0.5, 1.5))
        meshes.append((V, F))
    pos = rng.uniform(0.0, box, size=(n_bodies, 3))
    trees, rec, radii = _bodies(meshes, rng, pos, 1.0, pcg_seed(42, 54))
    s = Scene("c1", meshes, trees, rec, radii)
    s.pairs = broad_phase(rec, radii)
    return s


def scene_c2(seed=1002, n_bodies=4096, packing=0.30, vel_sigma=2.0):
    """4096 bodies: half subdivision-3 icospheres scaled U(0.5,1.5), half convex
    hulls of 64 points on ellipsoids with axes U(0.3,1.2); Poisson-disk
    placement of the bounding spheres at 30% packing; velocities N(0,2)."""
    rng = np.random.default_rng(seed)
    kinds = rng.permutation(np.arange(n_bodies) % 2)
    meshes = []
    for k in kinds:
        meshes.append(icosphere(3, rng.uniform(0.5, 1.5)) if k == 0 else ellipsoid_hull(rng))
    radii0 = np.array([np.max(np.linalg.norm(V, axis=1)) for V, _ in meshes])
    box = (np.sum(4.0 / 3.0 * np.pi * radii0 ** 3) / packing) ** (1.0 / 3.0)
    pos = poisson_disk(rng, radii0, box)
    trees, rec, radii = _bodies(meshes, rng, pos, vel_sigma, pcg_seed(42, 54))
    s = Scene("c2", meshes, trees, rec, radii)
    s.box = box
    s.pairs = broad_phase(rec, radii)
    return s


class FlatTrees:
    """Many trees in the flat layout d2_trees_upload takes: tree t owns node
    and bucket records offs[t]..offs[t+1].  Indexing gives (nodes, buckets)
    views, so it stands in for a list of trees."""

    def __init__(self, offs, nodes, buckets):
        self.offs, self.nodes, self.buckets = offs, nodes, buckets

    def __len__(self):
        return self.offs.size - 1

    def __getitem__(self, t):
        if isinstance(t, slice):
            return [self[i] for i in range(*t.indices(len(self)))]
        a, b = self.offs[t], self.offs[t + 1]
        return self.nodes[a:b], self.buckets[a:b]

    def __iter__(self):
        return (self[t] for t in range(len(self)))

    @staticmethod
    def from_list(trees):
        """Packs a list of trees, freeing each list entry as it is copied (the
        list then holds None), so the peak stays near one copy."""
        offs = np.zeros(len(trees) + 1, dtype=np.int32)
        offs[1:] = np.cumsum([t[0].shape[0] for t in trees])
        nodes = np.empty((int(offs[-1]), 8), dtype=np.int32)
        buckets = np.empty((int(offs[-1]), 88))
        for k in range(len(trees)):
            nodes[offs[k]:offs[k + 1]], buckets[offs[k]:offs[k + 1]] = trees[k]
            trees[k] = None
        return FlatTrees(offs, nodes, buckets)


def build_trees_flat(mesh_of, n, stream_base, threads=None, progress=None, first=None):
    """Trees of meshes mesh_of(0..n-1) on a thread pool, packed flat.  Mesh k
    uses its own PCG32 stream (42, stream_base + k), so the result does not
    depend on the thread count; ``first`` optionally prepends a prebuilt tree.
    The meshes are made on the fly and not kept.  progress(msg) is called
    about every 10% of the trees."""
    import os
    import time
    from concurrent.futures import ThreadPoolExecutor
    threads = threads or min(16, len(os.sched_getaffinity(0)))

    def one(k):
        V, F = mesh_of(k)
        return tree_build(V, F, pcg_seed(42, stream_base + k))

    trees = [first] if first is not None else []
    t0 = time.time()
    step = max(1, n // 10)
    with ThreadPoolExecutor(max_workers=threads) as ex:
        for k, t in enumerate(ex.map(one, range(n), chunksize=64)):
            trees.append(t)
            if progress and (k + 1) % step == 0:
                progress(f"trees {k + 1}/{n} in {time.time() - t0:.0f}s")
    return FlatTrees.from_list(trees)


def build_trees_parallel(meshes, stream_base, threads=None):
    """Trees of many meshes on a thread pool (the library call releases the
    GIL); mesh k uses its own PCG32 stream (42, stream_base + k), so the
    result does not depend on the thread count."""
    import os
    from concurrent.futures import ThreadPoolExecutor
    threads = threads or min(16, len(os.sched_getaffinity(0)))

    def one(k):
        V, F = meshes[k]
        return tree_build(V, F, pcg_seed(42, stream_base + k))

    with ThreadPoolExecutor(max_workers=threads) as ex:
        return list(ex.map(one, range(len(meshes)), chunksize=64))


def scene_c3(seed=1003, n_scenes=256, lo=16, hi=512, threads=None):
    """Batched workload: n_scenes independent scenes with round(exp(U(ln lo,
    ln hi))) bodies each; C1 meshes (subdivision-2 icospheres scaled
    U(0.5,1.5)), random rotations, positions U([0, side]^3) with side =
    10 (N/100)^(1/3) (the C1 density), velocities N(0,1), a random gravity
    direction per scene (|g| = 9.8; it acts between steps, not inside the
    contact-detection step).  Scenes share space but never share pairs: the
    broad phase runs per scene.  Trees use one PCG32 stream per body
    (42, 10000 + body) so they can be built in parallel."""
    rng = np.random.default_rng(seed)
    counts = np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), size=n_scenes))).astype(np.int64)
    V0, F0 = icosphere(2, 1.0)
    _, _, I0 = mass_properties(V0, F0)
    r0 = float(np.max(np.linalg.norm(V0, axis=1)))
    meshes, recs, radii, sob = [], [], [], []
    gravity = np.zeros((n_scenes, 3))
    for k, n in enumerate(counts):
        g = rng.normal(size=3)
        gravity[k] = 9.8 * g / np.linalg.norm(g)
        side = 10.0 * (n / 100.0) ** (1.0 / 3.0)
        for _ in range(n):
            sc = rng.uniform(0.5, 1.5)
            meshes.append((V0 * sc, F0))
            q = random_rotation(rng)
            pos = rng.uniform(0.0, side, size=3)
            v = rng.normal(0.0, 1.0, size=3)
            recs.append(body_record(q, pos, v, np.linalg.inv(I0 * sc ** 5)))
            radii.append(r0 * sc)
            sob.append(k)
    trees = build_trees_parallel(meshes, 10000, threads)
    rec = np.array(recs)
    radii = np.array(radii)
    sob = np.array(sob, dtype=np.int32)
    s = Scene("c3", meshes, trees, rec, radii, sob)
    s.gravity = gravity
    s.counts = counts
    first = np.concatenate([[0], np.cumsum(counts)])
    parts = []
    for k in range(n_scenes):
        p = broad_phase(rec[first[k]:first[k + 1]], radii[first[k]:first[k + 1]])
        parts.append(p + first[k])
    s.pairs = np.concatenate(parts).astype(np.int32) if parts else np.zeros((0, 2), dtype=np.int32)
    s.scene_first = first
    return s


def scene_c4(seed=1004, n_bodies=4096, threads=None):
    """Held-out evaluation scene of the learned gate: n_bodies C1 meshes
    (subdivision-2 icospheres scaled U(0.5,1.5)) at the C1 density (box side
    10 (n/100)^(1/3) m), random rotations, velocities N(0,1); trees from
    per-body PCG32 streams (42, 30000 + body)."""
    rng = np.random.default_rng(seed)
    V0, F0 = icosphere(2, 1.0)
    _, _, I0 = mass_properties(V0, F0)
    r0 = float(np.max(np.linalg.norm(V0, axis=1)))
    side = 10.0 * (n_bodies / 100.0) ** (1.0 / 3.0)
    meshes, recs, radii = [], [], []
    for _ in range(n_bodies):
        sc = rng.uniform(0.5, 1.5)
        meshes.append((V0 * sc, F0))
        recs.append(body_record(random_rotation(rng), rng.uniform(0.0, side, size=3), rng.normal(0.0, 1.0, size=3),
                                np.linalg.inv(I0 * sc ** 5)))
        radii.append(r0 * sc)
    trees = build_trees_parallel(meshes, 30000, threads)
    rec = np.array(recs)
    radii = np.array(radii)
    s = Scene("c4", meshes, trees, rec, radii)
    s.box = side
    s.pairs = broad_phase(rec, radii)
    return s


def funnel_mesh(r_top=20.0, r_bottom=4.0, height=40.0, n_theta=256, n_z=64):
    """Open cone frustum (axis z, outlet at z = 0), 2 n_theta n_z triangles,
    normals facing the axis."""
    th = np.linspace(0.0, 2.0 * np.pi, n_theta, endpoint=False)
    zs = np.linspace(0.0, height, n_z + 1)
    rs = r_bottom + (r_top - r_bottom) * zs / height
    V = np.array([[r * np.cos(t), r * np.sin(t), z] for z, r in zip(zs, rs) for t in th])
    F = []
    for i in range(n_z):
        for j in range(n_theta):
            a, b = i * n_theta + j, i * n_theta + (j + 1) % n_theta
            c, d = a + n_theta, b + n_theta
            F.append([a, c, b])
            F.append([b, c, d])
    return V, np.array(F, dtype=np.int32)


def funnel_distance(p, r_top=20.0, r_bottom=4.0, height=40.0):
    """Distance of points p (n, 3) from the funnel's cone surface (its segment
    in the (radius, z) half-plane)."""
    rr = np.hypot(p[:, 0], p[:, 1])
    a = np.array([r_bottom, 0.0])
    b = np.array([r_top, height])
    ab = b - a
    q = np.stack([rr, p[:, 2]], axis=1) - a
    t = np.clip(q @ ab / (ab @ ab), 0.0, 1.0)
    return np.linalg.norm(q - t[:, None] * ab, axis=1)


def scene_c5(seed=1005, n_bodies=65536, threads=None, density=0.1, subdiv=3, progress=None):
    """Large pile: n_bodies icospheres scaled U(0.5,1.5) -- subdivision 3
    (1280 triangles, 84 M in all at 65536 bodies) by default, subdivision 2
    (the 320 of C1) on request -- at the C1 number density (0.1 per m^3)
    inside a static 40 m funnel (top radius 20 m, outlet radius 4 m) and in
    the column of radius 20 m above it, falling at v ~ N(0,1) + (0,0,-3) m/s.
    Body 0 is the funnel (static, 32768 triangles).  A body pairs with the
    funnel when its bounding sphere comes within reach of the cone surface
    (what a triangle BVH broad phase decides); body pairs use the swept
    bounding-sphere test.  Trees use one PCG32 stream per mesh (42, 40000 +
    body), built in parallel and packed flat (FlatTrees); the meshes are not
    kept."""
    rng = np.random.default_rng(seed)
    V0, F0 = icosphere(subdiv, 1.0)
    _, _, I0 = mass_properties(V0, F0)
    r0 = float(np.max(np.linalg.norm(V0, axis=1)))
    funnel_v = np.pi * 40.0 / 3.0 * (20.0 ** 2 + 20.0 * 4.0 + 4.0 ** 2)
    col_h = max(0.0, (n_bodies / density - funnel_v) / (np.pi * 20.0 ** 2))
    pos = np.zeros((n_bodies, 3))
    k = 0
    while k < n_bodies:  # uniform in funnel + column (rejection sampling)
        m = (n_bodies - k) * 2
        c = np.stack([rng.uniform(-20, 20, m), rng.uniform(-20, 20, m), rng.uniform(0, 40 + col_h, m)], axis=1)
        rr = np.hypot(c[:, 0], c[:, 1])
        rmax = np.where(c[:, 2] < 40.0, 4.0 + 16.0 * c[:, 2] / 40.0, 20.0)
        c = c[rr < rmax - 0.3]
        take = min(len(c), n_bodies - k)
        pos[k:k + take] = c[:take]
        k += take
    FV, FF = funnel_mesh()
    recs = [body_record(np.array([1.0, 0.0, 0.0, 0.0]), np.zeros(3), np.zeros(3), np.zeros((3, 3)), static=True)]
    radii = [float(np.max(np.linalg.norm(FV, axis=1)))]
    scale = np.zeros(n_bodies + 1)
    for i in range(n_bodies):
        sc = rng.uniform(0.5, 1.5)
        scale[i + 1] = sc
        v = rng.normal(0.0, 1.0, size=3) + np.array([0.0, 0.0, -3.0])
        recs.append(body_record(random_rotation(rng), pos[i], v, np.linalg.inv(I0 * sc ** 5)))
        radii.append(r0 * sc)
    trees = build_trees_flat(lambda k: (FV, FF) if k == 0 else (V0 * scale[k], F0), n_bodies + 1, 40000, threads,
                             progress)
    rec = np.array(recs)
    radii = np.array(radii)
    s = Scene("c5", [], trees, rec, radii)
    s._n_tris = int(FF.shape[0] + n_bodies * F0.shape[0])
    s.subdiv = subdiv
    s.pairs = c5_pairs(rec, radii)
    return s


def c5_pairs(rec, radii, geo_eps=GEO_EPS):
    """Body-body pairs by the swept-sphere test, plus (0, body) for bodies
    within reach of the funnel wall."""
    body = broad_phase(rec[1:], radii[1:], geo_eps) + 1
    reach = radii[1:] + 2 * geo_eps + np.linalg.norm(rec[1:, 7:10], axis=1) * rec[1:, 39]
    near = np.nonzero(funnel_distance(rec[1:, 4:7]) < reach)[0] + 1
    wall = np.stack([np.zeros(near.size, dtype=np.int64), near], axis=1)
    return np.concatenate([wall, body]).astype(np.int32)


def advance_gravity(records, g=(0.0, 0.0, -9.8), dt=DT):
    """Ballistic step of the dynamic bodies under gravity (no contact response)."""
    r = records.copy()
    dyn = r[:, 38] == 0
    r[dyn, 7:10] += dt * np.asarray(g)
    r[dyn, 4:7] += dt * r[dyn, 7:10]
    r[:, 13] += dt
    r[:, 14:28] = r[:, 0:14]
    return r


def advance(records, dt=DT):
    """Ballistic drift of the current state by dt (no contact response: that is
    the solver's job, outside this path)."""
    r = records.copy()
    r[:, 4:7] += dt * r[:, 7:10]
    r[:, 13] += dt
    r[:, 14:28] = r[:, 0:14]
    return r
