"""Broad-phase scenes shared by the CPU oracle tests and the GPU parity tests.

Meshes from the reference's own test fixtures (common/primitive_geo.cpp:6-31
cube, :76-89 plane), body defaults of Particle(mesh, 1.0, 1.0, 0.01)
(model/particle.cpp:7-23: geo_eps 0.01, last_time_step_size 1.0, identity
state at time 0), and seeded random scenes for the brute-force comparisons.
"""
import numpy as np

from delta2_amd import scenes

CUBE_V = np.array([[1.0, 1.0, 1.0], [1.0, 1.0, -1.0], [1.0, -1.0, 1.0], [1.0, -1.0, -1.0],
                   [-1.0, 1.0, 1.0], [-1.0, 1.0, -1.0], [-1.0, -1.0, 1.0], [-1.0, -1.0, -1.0]])
CUBE_F = np.array([[0, 2, 1], [1, 2, 3], [2, 6, 3], [3, 6, 7], [4, 5, 6], [5, 7, 6], [4, 0, 1], [4, 1, 5],
                   [0, 4, 6], [0, 6, 2], [3, 7, 1], [7, 5, 1]], dtype=np.int32)


def plane_mesh(width):
    h = width / 2.0
    V = np.array([[h, h, 0.0], [h, -h, 0.0], [-h, h, 0.0], [-h, -h, 0.0]])
    F = np.array([[0, 2, 1], [1, 2, 3]], dtype=np.int32)
    return V, F


def particle(pos=(0.0, 0.0, 0.0), vel=(0.0, 0.0, 0.0), static=False):
    """Particle(mesh, 1.0, 1.0, 0.01) with its translation and velocity set."""
    return scenes.body_record([1.0, 0.0, 0.0, 0.0], pos, vel, np.eye(3), geo_eps=0.01, static=static, last_step=1.0)


class BpScene:
    """records (n, 40), last states (n, 14), per-body mesh bounds, groups."""

    def __init__(self, records, aabb6, radius, last=None, groups=None, mesh_of_body=None):
        self.records = np.asarray(records, dtype=np.float64).reshape(-1, 40)
        self.aabb6 = np.asarray(aabb6, dtype=np.float32).reshape(-1, 6)
        self.radius = np.asarray(radius, dtype=np.float64)
        self.last = None if last is None else np.asarray(last, dtype=np.float64).reshape(-1, 14)
        self.groups = None if groups is None else np.asarray(groups, dtype=np.int32)
        self.mesh_of_body = mesh_of_body

    @property
    def n(self):
        return self.records.shape[0]


def reference_cases(bounds):
    """The scenes of collision_detection/test_broad_phase.cpp:6-120 as
    (name, scene, mode 0 step t, expected pair count).  bounds(V) gives a
    mesh's (aabb6, radius)."""
    cb, cr = bounds(CUBE_V)
    pb, pr = bounds(plane_mesh(10)[0])
    out = []
    # :6-27 ten stationary cubes 2 m apart
    recs = [particle((2.0 * i, 0.0, 0.0)) for i in range(10)]
    out.append(("stationary", BpScene(recs, [cb] * 10, [cr] * 10), 1.0, 9))
    # :30-61 two cubes approaching
    recs = [particle((0.0, 0.0, 0.0), (1.0, 0.0, 0.0)), particle((4.0, 0.0, 0.0), (-1.0, 0.0, 0.0))]
    s = BpScene(recs, [cb] * 2, [cr] * 2)
    out += [("moving", s, 0.1, 0), ("moving", s, 0.5, 1)]
    # :63-110 a cube falling onto a static plane, then onto two planes
    recs = [particle(static=True), particle((0.0, 0.0, 2.0), (0.0, 0.0, -1.0))]
    s = BpScene(recs, [pb, cb], [pr, cr])
    out += [("static", s, 0.1, 0), ("static", s, 1.0, 1)]
    recs = recs + [particle(static=True)]
    out.append(("static2", BpScene(recs, [pb, cb, pb], [pr, cr, pr]), 1.0, 2))
    return out


def random_scene(seed, n=300, box=12.0, n_static=3, n_groups=1, uneven_times=False, vel_sigma=2.0):
    """Icospheres (subdivision 1, radius U(0.4, 1.2)) at uniform positions
    with N(0, vel_sigma) velocities and last steps U(dt/4, dt), plus n_static
    static square plates (width 0.8 box).  Returns (meshes, records, last
    states, groups); meshes[i] = (V, F) of body i.  With uneven_times every
    non-static body's current time is 1 + U(0, dt/4) and its last state lies
    one last step earlier (so any other body's time falls inside it)."""
    rng = np.random.default_rng(seed)
    V0, F0 = scenes.icosphere(1, 1.0)
    PV, PF = plane_mesh(box * 0.8)
    static_ids = set(rng.choice(n, size=min(n_static, n), replace=False).tolist()) if n_static else set()
    meshes, recs, last = [], [], []
    for i in range(n):
        static = i in static_ids
        if static:
            V, F = PV, PF
            vel = np.zeros(3)
        else:
            V, F = V0 * rng.uniform(0.4, 1.2), F0
            vel = rng.normal(0.0, vel_sigma, size=3)
        pos = rng.uniform(0.0, box, size=3)
        r = scenes.body_record(scenes.random_rotation(rng), pos, vel, np.eye(3), geo_eps=0.01, static=static,
                               last_step=float(rng.uniform(0.25, 1.0) * scenes.DT))
        r[13] = 1.0 + (float(rng.uniform(0.0, 0.25 * scenes.DT)) if uneven_times and not static else 0.0)
        # future: the translation advanced by the last step (what the broad phase reads of it)
        r[14:28] = r[0:14]
        r[18:21] = r[4:7] + r[39] * r[7:10]
        r[27] = r[13] + r[39]
        ls = r[0:14].copy()
        if uneven_times and not static:
            ls[0:4] = scenes.random_rotation(rng)
            ls[4:7] = r[4:7] - r[39] * r[7:10] + rng.normal(0.0, 0.01, size=3)
            ls[7:10] = r[7:10] + rng.normal(0.0, 0.1, size=3)
            ls[10:13] = rng.normal(0.0, 0.1, size=3)
            ls[13] = r[13] - r[39]
        meshes.append((V, F))
        recs.append(r)
        last.append(ls)
    groups = rng.integers(0, n_groups, size=n).astype(np.int32) if n_groups > 1 else None
    return meshes, np.array(recs), np.array(last), groups
