"""Shared inputs for the parity tests: scenes plus the oracle's view of them."""
import functools

import numpy as np

import d2o
from delta2_amd import scenes


def with_future(records, step):
    """Body records whose future state is current.extrapolate(step) (oracle)."""
    r = records.copy()
    for i in range(r.shape[0]):
        r[i, 14:28] = d2o.extrapolate(r[i, 0:14], r[i, 28:37], step[i] if np.ndim(step) else step)
    return r


@functools.lru_cache(maxsize=None)
def c1_scene():
    return scenes.scene_c1()


@functools.lru_cache(maxsize=None)
def dense_scene(seed=7, n=60, box=5.0, spin=2.0):
    """C1 meshes packed into a small box with spin, so pairs interpenetrate:
    exercises dist == 0, toc == 0, many leaf contacts and filter ties."""
    rng = np.random.default_rng(seed)
    meshes = [scenes.icosphere(2, rng.uniform(0.5, 1.5)) for _ in range(n)]
    pos = rng.uniform(0.0, box, size=(n, 3))
    trees, rec, radii = scenes._bodies(meshes, rng, pos, 1.0, _seed())
    rec[:, 10:13] = rng.normal(0.0, spin, size=(n, 3))  # angular momentum
    rec[:, 14:28] = rec[:, 0:14]
    s = scenes.Scene("dense", meshes, trees, rec, radii)
    s.pairs = scenes.broad_phase(rec, radii)
    return s


def _seed():
    from delta2_amd import pcg_seed
    return pcg_seed(42, 54)


@functools.lru_cache(maxsize=None)
def hull_scene(seed=11, n=40, box=4.0):
    """Convex hulls (thin triangles, off-centre geometry) packed densely."""
    rng = np.random.default_rng(seed)
    meshes = [scenes.ellipsoid_hull(rng) for _ in range(n)]
    pos = rng.uniform(0.0, box, size=(n, 3))
    trees, rec, radii = scenes._bodies(meshes, rng, pos, 2.0, _seed())
    s = scenes.Scene("hulls", meshes, trees, rec, radii)
    s.pairs = scenes.broad_phase(rec, radii)
    return s


def oracle_scene(s, records):
    return d2o.Scene(s.trees, s.tree_of_body, records)


@functools.lru_cache(maxsize=None)
def sparse_scene(seed=1202, n=256):
    """The 4096-body workload's generator at 256 bodies (half subdivision-3
    icospheres, half ellipsoid hulls, Poisson-disk placement at 30% packing):
    many broad-phase pairs are separated, some touch."""
    return scenes.scene_c2(seed=seed, n_bodies=n)
