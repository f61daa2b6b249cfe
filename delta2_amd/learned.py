"""Learned pair gate: training data from plain CCD runs, device training,
weights on disk, evaluation against the plain path.

Labels per pair come from the plain CCD + TOI -> step rule (the library's
exact path): y0 = the pair's step scale, y1 = the smallest toc of its CCD
contacts (FREE_T = 1.25 when it has none).  Features are the library's
d2_pair_features.  Training runs the library's HIP forward/backward/Adam
kernels (d2_mlp_train); inference is its MFMA kernel.  The gate proposes a
pair as contact-free when y0 >= S_GATE and y1 >= T_GATE; a proposal only
decides whether the pair's separation certificate is tried, never a result.

Training episodes are C1-style scenes (subdivision-2 icospheres scaled
U(0.5, 1.5), velocities N(0, 1), dt = 1/60): a pool of bodies whose trees are
built once, placed anew in every episode (box side 10 m times U(1, 2.2) so
both touching and separated pairs occur) and advanced ballistically under
gravity for a few steps, with the CCD labels logged at every step.
"""
import os

import numpy as np

from . import D2_FP32, Context, scenes

FREE_T = 1.25
T_GATE = 1.1
S_GATE = 0.99
LW = (1.0, 1.0)
DELTA = 0.1
N_IN, N_H, N_OUT = 24, 64, 2
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "pair_gate_mlp.npy")


def init_params(seed=0):
    """He-normal weights and zero biases in the library's parameter layout."""
    rng = np.random.default_rng(seed)
    parts = [rng.normal(0, np.sqrt(2.0 / N_IN), (N_IN, N_H)), np.zeros(N_H),
             rng.normal(0, np.sqrt(2.0 / N_H), (N_H, N_H)), np.zeros(N_H),
             rng.normal(0, np.sqrt(1.0 / N_H), (N_H, N_OUT)), np.zeros(N_OUT)]
    return np.concatenate([p.ravel() for p in parts]).astype(np.float32)


def labels(result, n_pairs):
    """(scale, min toc or FREE_T) per pair from a CCD result with per-pair values."""
    t = np.full(n_pairs, FREE_T)
    if result.n:
        np.minimum.at(t, result.ints[:, 0], result.reals[:, 10])
    return np.stack([result.per_pair[:, 0], t], axis=1).astype(np.float32)


class TrainingBodies:
    """A pool of C1-style bodies (meshes, trees, inertia) in n_scenes groups."""

    def __init__(self, n_scenes=32, per_scene=100, seed=2004, threads=None):
        rng = np.random.default_rng(seed)
        V0, F0 = scenes.icosphere(2, 1.0)
        _, _, I0 = scenes.mass_properties(V0, F0)
        r0 = float(np.max(np.linalg.norm(V0, axis=1)))
        self.n_scenes, self.per_scene = n_scenes, per_scene
        n = n_scenes * per_scene
        self.scale = rng.uniform(0.5, 1.5, size=n)
        meshes = [(V0 * s, F0) for s in self.scale]
        self.trees = scenes.build_trees_parallel(meshes, 20000, threads)
        self.inv = np.stack([np.linalg.inv(I0 * s ** 5) for s in self.scale])
        self.radii = r0 * self.scale
        self.tree_of_body = np.arange(n, dtype=np.int32)
        self.scene_of_body = np.repeat(np.arange(n_scenes), per_scene).astype(np.int32)

    @property
    def n_bodies(self):
        return self.scale.size

    def place(self, rng):
        """Fresh random placement of every scene: records (n, 40)."""
        rec = np.zeros((self.n_bodies, 40))
        side = 10.0 * rng.uniform(1.0, 2.2, size=self.n_scenes)
        for i in range(self.n_bodies):
            k = self.scene_of_body[i]
            rec[i] = scenes.body_record(scenes.random_rotation(rng), rng.uniform(0.0, side[k], size=3),
                                        rng.normal(0.0, 1.0, size=3), self.inv[i])
        return rec

    def pairs(self, rec):
        out = []
        for k in range(self.n_scenes):
            b0 = k * self.per_scene
            p = scenes.broad_phase(rec[b0:b0 + self.per_scene], self.radii[b0:b0 + self.per_scene])
            out.append(p + b0)
        return np.concatenate(out).astype(np.int32)


def advance(rec, gravity, dt=scenes.DT):
    """Ballistic step under gravity (no contact response)."""
    r = rec.copy()
    r[:, 7:10] += dt * gravity
    r[:, 4:7] += dt * r[:, 7:10]
    r[:, 13] += dt
    r[:, 14:28] = r[:, 0:14]
    return r


def make_dataset(ctx, pool, n_target, seed=0, steps_per_episode=4, log=None):
    """Features and labels of about n_target pairs from plain CCD runs."""
    rng = np.random.default_rng(seed)
    ctx.upload_trees(pool.trees)
    Xs, Ts, total = [], [], 0
    while total < n_target:
        rec = pool.place(rng)
        g = rng.normal(size=3)
        g = 9.8 * g / np.linalg.norm(g)
        for _ in range(steps_per_episode):
            pairs = pool.pairs(rec)
            if pairs.shape[0]:
                ctx.upload_bodies(pool.tree_of_body, rec)
                mt = np.minimum(rec[pairs[:, 0], 39], rec[pairs[:, 1], 39])
                ctx.upload_pairs(pairs, mt)
                ctx.ccd_run()
                Ts.append(labels(ctx.result(with_pairs=True), pairs.shape[0]))
                Xs.append(ctx.pair_features())
                total += pairs.shape[0]
            rec = advance(rec, g)
        if log:
            log(f"[learned] dataset {total}/{n_target} pairs")
    X = np.concatenate(Xs)[:n_target]
    T = np.concatenate(Ts)[:n_target]
    return X, T


def train(ctx, X, T, epochs=12, batch=4096, lr=2e-3, seed=0, params=None):
    """Trains on the device from He-initialised (or given) weights; returns
    (params, per-epoch mean loss)."""
    ctx.mlp_set_params(init_params(seed) if params is None else params)
    losses = []
    # step decay: lr, lr/4, lr/16 over the three thirds of the epochs
    for k, f in enumerate((1.0, 0.25, 0.0625)):
        e = max(1, epochs // 3 + (1 if k < epochs % 3 else 0))
        losses.extend(ctx.mlp_train(X, T, e, batch=batch, lr=lr * f, lw=LW, delta=DELTA, seed=seed + 17 * k))
    return ctx.mlp_get_params(), np.array(losses)


def gate_stats(pred, T):
    """Confusion of the gate's proposals against the labels' contact-free pairs."""
    prop = (pred[:, 0] >= S_GATE) & (pred[:, 1] >= T_GATE)
    free = T[:, 1] >= FREE_T - 1e-6
    return {"proposed": float(prop.mean()), "free": float(free.mean()),
            "precision": float((prop & free).sum() / max(prop.sum(), 1)),
            "recall": float((prop & free).sum() / max(free.sum(), 1))}


def save_weights(params, path=WEIGHTS):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.save(path, np.asarray(params, dtype=np.float32))


def load_weights(path=WEIGHTS):
    return np.load(path) if os.path.exists(path) else None


def main():
    import argparse
    import sys
    import time
    ap = argparse.ArgumentParser(description="train the learned pair gate on plain CCD labels and save its weights")
    ap.add_argument("--pairs", type=int, default=2_000_000)
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--out", default=WEIGHTS)
    args = ap.parse_args()
    log = lambda *a: print(*a, file=sys.stderr, flush=True)  # noqa: E731
    t0 = time.time()
    pool = TrainingBodies()
    log(f"[learned] body pool {pool.n_bodies} in {time.time() - t0:.1f}s")
    with Context(0, D2_FP32) as ctx:
        X, T = make_dataset(ctx, pool, args.pairs, seed=1, log=log)
        Xv, Tv = make_dataset(ctx, pool, max(20000, args.pairs // 20), seed=99)
        t1 = time.time()
        params, losses = train(ctx, X, T, epochs=args.epochs)
        log(f"[learned] trained {args.epochs} epochs on {X.shape[0]} pairs in {time.time() - t1:.1f}s; "
            f"loss {losses[0]:.4f} -> {losses[-1]:.4f}")
        val_loss, _ = ctx.mlp_loss_grad(Xv, Tv, LW, DELTA)
        stats = gate_stats(ctx.mlp_infer(Xv), Tv)
        log(f"[learned] validation loss {val_loss:.4f}, gate {stats}")
    save_weights(params, args.out)
    log(f"[learned] weights -> {args.out}")


if __name__ == "__main__":
    main()
