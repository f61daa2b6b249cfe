"""Benchmark: narrow+CCD pair queries/s on one MI355X.

A step is one pass of the contact-detection path over the step's broad-phase
pair list with its inputs resident in HBM (d2_step_run):
  CCD (compareTreesFullContinuous over [0, min(last steps)]) + the TOI -> step
  rule, the clamped cluster steps, the future states at those steps, then the
  discrete narrow phase (compareTreesFull) + filterContacts,
for every pair.  value = pairs summed over the timed steps / the summed wall
time of their d2_step_run calls (each bracketed by stream synchronisation).

Default workload (the largest single-GPU configuration): the large pile --
65 536 subdivision-3 icospheres (83.9 M triangles) in and above a static 40 m
funnel, falling; each step's pair list and clusters come from the device
broad phase and cluster separation over that step's body states, as the
reference's cluster stepper forms them (between steps the bodies move
ballistically: contact response is the solver's, outside this path).  The
learned + FP32 form and the plain FP64 path run the same steps; their
results are compared step by step.  The line also carries:
  roofline      the dominant kernel's algorithmic bytes per launch (the
                (body, node) records its levels test, counted on the device)
                over its GPU-stamped launch time, against the HBM peak; and the
                whole step's algorithmic bytes over the step time;
  cpu_baseline  the oracle (the reference path restated in C++) on the host
                cores over a sample of the same step's pairs;
  c2            the 4096-body workload's figures (plain FP64 vs this form).

Prints one JSON line.  ``--gpus N`` with N > 1 exits: the path runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
METRIC = "narrow+CCD pair queries/s, 1 GPU; learned vs plain speedup at equal TOI error"
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")
P64 = 8  # byte width of the reals the traversal kernels read (FP64 buckets and poses)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# --- algorithmic bytes (per node record tested, per body, per pair, per hit) ----
def node_bytes(c, P=P64):
    """Bytes of the records behind touch counts c = (inner nodes, leaf nodes,
    leaf triangles, bodies): a node's 16 B record, an inner bucket's 8 x 11
    reals, a leaf's n x 9 reals, a body's 40 reals."""
    return int(c[0]) * (16 + 88 * P) + int(c[1]) * 16 + int(c[2]) * 9 * P + int(c[3]) * 40 * P


def level_bytes(per_level, phase, P=P64):
    """Summed over a phase's levels: what its per-level kernels must read."""
    return sum(node_bytes(per_level[phase, lv], P) for lv in range(per_level.shape[1]))


def step_alg_bytes(run_counts, n_pairs, raw_hits, P=P64):
    """One step's algorithmic bytes: the distinct (body, node) records of both
    traversals, 16 B per pair (ids in, step scale out), 11 reals + 16 B per
    raw hit."""
    return node_bytes(run_counts, P) + 16 * n_pairs + raw_hits * (11 * P + 16)


# phase whose levels each traversal-cull kernel runs (0 discrete, 1 CCD)
KERNEL_PHASE = {"tri_mask": 0, "tri_full": 0, "soup_cull": 1}


def roofline_of(stats, per_level, traffic_key):
    """The dominant kernel among the per-level culls: its bytes per launch
    over its average GPU-stamped launch time."""
    dom = max((k for k in KERNEL_PHASE if stats.get(k, (0, 0.0, 0))[0]), key=lambda k: stats[k][1])
    n, ms, _ = stats[dom]
    avg_s = ms / n / 1e3
    per_launch = level_bytes(per_level, KERNEL_PHASE[dom]) / n
    achieved = per_launch / avg_s / 1e9
    traffic = None
    if os.path.exists(TRAFFIC_FILE):  # measured HBM bytes per dispatch (rocprofv3 FETCH_SIZE + WRITE_SIZE, calibrated)
        with open(TRAFFIC_FILE) as f:
            t = json.load(f).get(traffic_key, {}).get(dom)
        traffic = t.get("bytes_per_dispatch") if t else None
    return {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "avg_launch_us": avg_s * 1e6,
            "alg_bytes_per_launch": per_launch,
            "bytes_model": "per level: distinct (body, node) records the level tests (device count) x "
                           "[16 + 88P inner | 16 + 9nP leaf] + 40P per body, P = 8"}


def kernel_table(ctx, tag):
    stats = ctx.kernel_stats()
    for k, (n, ms, items) in sorted(stats.items(), key=lambda kv: -kv[1][1]):
        if n:
            log(f"[{tag}]   {k:12s} launches {n:6d}  total {ms:9.3f} ms  avg {ms / n * 1e3:9.2f} us  items {items:.0f}")
    return stats


# --- the CPU baseline -------------------------------------------------------------
def cpu_baseline(trees, tree_of_body, records, pairs, max_time, budget_s=12.0):
    """The oracle (C++ restatement of the reference path) on the host cores,
    on a bounded random sample of the pairs: CCD + step rule, then discrete +
    filter per pair (the discrete phase at the records' future states).
    Only the sampled pairs' bodies and trees are handed to the oracle."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import d2o
    d2o.lib()
    rng = np.random.default_rng(0)
    order = rng.permutation(pairs.shape[0])

    def sub(sel):
        bodies, inv = np.unique(pairs[sel], return_inverse=True)
        tids = np.unique(tree_of_body[bodies])
        tmap = {int(t): i for i, t in enumerate(tids)}
        sc = d2o.Scene([trees[int(t)] for t in tids], [tmap[int(tree_of_body[b])] for b in bodies], records[bodies])
        return sc, inv.reshape(-1, 2).astype(np.int32)

    n = min(16, pairs.shape[0])
    sc, p = sub(order[:n])
    t1, _ = sc.time_pass(p, max_time[order[:n]], threads=1, repeats=1)
    per_pair = max(t1 / n, 1e-6)
    n = int(min(pairs.shape[0], max(n, budget_s / 3.0 / per_pair)))
    sel = order[:n]
    sc, p = sub(sel)
    t1, _ = sc.time_pass(p, max_time[sel], threads=1, repeats=3)
    cores = len(os.sched_getaffinity(0))
    tn, _ = sc.time_pass(p, max_time[sel], threads=cores, repeats=3)
    return {"value": n / tn, "unit": "pair queries/s", "cores": cores, "kind": "port", "value_1_thread": n / t1,
            "sample": f"{n} of {pairs.shape[0]} pairs of the measured step drawn uniformly (seed 0): CCD + step rule, "
                      f"then discrete + filter at the step's future states; median of 3 passes; oracle built with "
                      f"g++ -O3 -ffp-contract=off"}


def copy_calibration(mib):
    """A plain device-to-device copy of known size through the HIP runtime, so
    FETCH_SIZE / WRITE_SIZE counters can be scaled against exact byte counts."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = mib << 20
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(src), ctypes.c_size_t(n)) == 0
    assert hip.hipMalloc(ctypes.byref(dst), ctypes.c_size_t(n)) == 0
    assert hip.hipMemset(src, 1, ctypes.c_size_t(n)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(dst, src, ctypes.c_size_t(n), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)
    print(json.dumps({"copy_calibration_bytes": n, "copies": 3}))


def gate_weights():
    """The committed gate weights, or a quick model trained on the device."""
    from delta2_amd import learned
    w = learned.load_weights()
    if w is not None:
        return w, "delta2_amd/weights/pair_gate_mlp.npy"
    import delta2_amd as d2
    with d2.Context(0, d2.D2_FP32) as c:
        pool = learned.TrainingBodies(n_scenes=8)
        X, T = learned.make_dataset(c, pool, 200_000, seed=1)
        w, _ = learned.train(c, X, T, epochs=9)
    return w, "trained in this run on 200k pairs"


def step_results(ctx):
    r = ctx.result(with_pairs=True)
    return r.offsets.copy(), r.ints.copy(), r.reals.copy(), r.per_pair.copy()


def same_results(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


def tune_step(ctx, learned, budget_s, tag):
    """Coordinate descent over the launch shapes (d2_set_tune) -- the
    candidates of the library's d2_autotune, one sweep -- timing one
    d2_step_run per candidate on the resident inputs and logging each, until
    the time budget is spent.  Leaves the best shape set."""
    from delta2_amd import scenes
    t_start = time.time()
    best = ctx.get_tune()
    u = max(1, best[0] // 4)  # the default shape is (4, 8, 16) blocks per CU

    def run(shape):
        ctx.set_tune(*shape)
        ctx.sync()
        t0 = time.perf_counter()
        ctx.step_run(scenes.DT, True)
        ctx.sync()
        return 1e3 * (time.perf_counter() - t0)

    best_ms = run(best)
    log(f"[{tag}] tune {best}: {best_ms:.1f} ms")
    dims = [[u, 2 * u, 4 * u, 8 * u, 16 * u], [2 * u, 4 * u, 8 * u, 16 * u, 32 * u], [4 * u, 16 * u, 64 * u]]
    if learned:
        dims.append([0, 1, 2, 3])  # phase mask of the gate; 0 = it gates no phase
    for d, vals in enumerate(dims):
        for v in vals:
            if time.time() - t_start > budget_s:
                log(f"[{tag}] tune budget of {budget_s:.0f}s spent")
                ctx.set_tune(*best)
                return best, best_ms
            shape = list(best)
            if shape[d] == v or v > (4096 if d == 0 else 16384):
                continue
            shape[d] = v
            ms = run(tuple(shape))
            log(f"[{tag}] tune {tuple(shape)}: {ms:.1f} ms")
            if ms < best_ms:
                best, best_ms = tuple(shape), ms
    ctx.set_tune(*best)
    return best, best_ms


# --- the large pile (default) ------------------------------------------------------
class PileStepper:
    """One context stepping the pile: per step, the body states are uploaded,
    projected as the cluster stepper does before its broad phase
    (broad_phase_embree_cluster.cpp:126-132), then broad phase and cluster
    separation run on the device, then the timed d2_step_run."""

    def __init__(self, sc, prec, weights=None, prune=True):
        import delta2_amd as d2
        self.sc = sc
        self.ctx = d2.Context(0, prec)
        self.ctx.set_pruning(prune)
        t0 = time.time()
        self.ctx.upload_trees(sc.trees)
        self.upload_s = time.time() - t0
        if weights is not None:
            self.ctx.mlp_set_params(weights)
            self.ctx.set_learned(True)
        self.prep_ms = []

    def prepare(self, rec):
        from delta2_amd import scenes
        c = self.ctx
        c.upload_bodies(self.sc.tree_of_body, rec)
        c.project_future(np.minimum(2.0 * rec[:, 39], scenes.DT))
        c.sync()
        t0 = time.perf_counter()
        n = c.broad_phase(mode=1)
        t1 = time.perf_counter()
        c.build_clusters()
        c.sync()
        self.prep_ms.append((1e3 * (t1 - t0), 1e3 * (time.perf_counter() - t1)))
        return n

    def step(self, rec):
        """(pairs, seconds of d2_step_run)."""
        from delta2_amd import scenes
        n = self.prepare(rec)
        self.ctx.sync()
        t0 = time.perf_counter()
        self.ctx.step_run(scenes.DT, True)
        self.ctx.sync()
        return n, time.perf_counter() - t0

    def results(self):
        return step_results(self.ctx) + (self.ctx.download_bodies()[:, 14:28].copy(),)

    def close(self):
        self.ctx.close()


def pile_states(sc, n):
    """The pile's body states for n steps: each step advances the previous one
    ballistically under gravity."""
    from delta2_amd import scenes
    out = [sc.records]
    for _ in range(n - 1):
        out.append(scenes.advance_gravity(out[-1]))
    return out


def run_pile(args):
    import delta2_amd as d2
    from delta2_amd import scenes
    t_start = time.time()
    name = f"c5_s{args.subdiv}_{args.bodies}"
    sc = scenes.cached(name, lambda: scenes.scene_c5(n_bodies=args.bodies, subdiv=args.subdiv,
                                                       progress=lambda m: log(f"[c5] {m}")))
    build_s = time.time() - t_start
    log(f"[c5] {sc.n_bodies} bodies, {sc.n_triangles()} triangles, {int(sc.trees.offs[-1])} tree nodes, "
        f"built in {build_s:.1f}s")
    weights, wsrc = gate_weights() if args.learned == "on" else (None, None)
    prec = d2.D2_FP32 if args.precision == "fp32" else d2.D2_FP64
    W, K = args.warmup, args.steps
    n_plain = K if args.plain_steps <= 0 else min(K, args.plain_steps)
    states = pile_states(sc, W + K)

    # this form: warm-up steps (the first sizes buffers, then the autotuner
    # picks launch shapes on its resident inputs), then the K timed steps
    fast = PileStepper(sc, prec, weights)
    log(f"[c5] trees uploaded in {fast.upload_s:.1f}s")
    n, s = fast.step(states[0])
    log(f"[c5] first step (sizes the buffers): {n} pairs, {1e3 * s:.1f} ms")
    if not args.no_autotune:
        shape, ms = tune_step(fast.ctx, weights is not None, args.tune_budget, "c5")
        log(f"[c5] tuned {shape}: {ms:.1f} ms")
    tuned = fast.ctx.get_tune()  # the gate's phase mask is its last entry (0: it gates no phase)
    for k in range(1, W):
        n, s = fast.step(states[k])
        log(f"[c5] warm-up step {k}: {n} pairs, {1e3 * s:.1f} ms")
    pairs_t, secs, ref = [], [], []
    for k in range(W, W + K):
        n, s = fast.step(states[k])
        pairs_t.append(n)
        secs.append(s)
        if k - W < n_plain and not args.no_plain:
            ref.append(fast.results())
        log(f"[c5] step {k - W}: {n} pairs, {1e3 * s:.1f} ms")
    value = sum(pairs_t) / sum(secs)
    ms_per_step = 1e3 * sum(secs) / K
    gate = fast.ctx.gate_counts() if weights is not None else None

    # what each part of this form buys: the last timed state stepped again
    # with it off (same results); the learned gate in each phase mask
    def timed_last():
        fast.step(states[W + K - 1])  # sizes the buffers for this form
        return 1e3 * fast.step(states[W + K - 1])[1]

    def ab(label, off, on):
        off()
        s_off = timed_last()
        on()
        s_on = timed_last()
        log(f"[c5] {label}: {s_on:.1f} ms with, {s_off:.1f} ms without")
        return {"ms_with": s_on, "ms_without": s_off}
    parts = {}
    if weights is not None:
        by = {}
        for name, ph in (("off", 0), ("ccd", 1), ("discrete", 2), ("both", 3)):
            fast.ctx.set_learned(ph != 0, phases=ph)
            by[name] = timed_last()
        fast.ctx.set_learned(True)
        fast.ctx.set_tune(*tuned)  # back to the tuned phase mask
        log(f"[c5] learned gate by phases (ms): {by}, tuned mask {tuned[3]}")
        parts["learned_gate"] = {"ms_by_phases": by, "tuned_phases": tuned[3]}
    parts["subtree_pruning"] = ab("subtree pruning", lambda: fast.ctx.set_pruning(False),
                                  lambda: fast.ctx.set_pruning(True))

    # per-kernel times (one eager step with HIP events around every kernel)
    # and the (body, node) records its levels test, on the last timed state
    last = states[W + K - 1]
    fast.prepare(last)
    fast.ctx.reset_stats()
    fast.ctx.set_profiling(True)
    fast.ctx.step_run(scenes.DT, True)
    fast.ctx.sync()
    fast.ctx.set_profiling(False)
    stats = kernel_table(fast.ctx, "c5")
    fast.ctx.set_touch_stats(True)
    fast.ctx.step_run(scenes.DT, True)
    fast.ctx.set_touch_stats(False)
    per_level, _ = fast.ctx.touch_stats()
    log("[c5] accounting step done")
    tag = f"c5/{args.precision}/{'learned' if weights is not None else 'plain'}"
    roofline = roofline_of(stats, per_level, tag)
    log(f"[c5] roofline {roofline['kernel']}: {roofline['achieved']:.0f} GB/s ({100 * roofline['frac']:.1f}% of HBM)")
    # end to end: host body states in, contacts and per-pair CCD values out
    t0 = time.perf_counter()
    fast.prepare(last)
    fast.ctx.step_run(scenes.DT, True)
    fast.ctx.result(with_pairs=True)
    e2e_ms = 1e3 * (time.perf_counter() - t0)
    log(f"[c5] end to end: {e2e_ms:.1f} ms")
    # the CPU baseline's inputs: the first timed step's pairs and future states
    fast.prepare(states[W])
    cpu_pairs, cpu_mt = fast.ctx.download_pairs()
    fast.ctx.step_run(scenes.DT, True)
    cpu_rec = fast.ctx.download_bodies()
    fast_prep = np.array(fast.prep_ms)
    fast.close()

    # the plain FP64 path on the same states
    plain_out, identical = None, None
    step_bytes = None
    if not args.no_plain:
        plain = PileStepper(sc, d2.D2_FP64, prune=False)  # the reference's traversal, every node pair
        log(f"[c5] plain fp64: trees uploaded in {plain.upload_s:.1f}s")
        n, s = plain.step(states[0])
        log(f"[c5] plain fp64 first step (sizes the buffers): {1e3 * s:.1f} ms")
        psecs, ppairs, identical = [], [], True
        for k in range(W, W + n_plain):
            n, s = plain.step(states[k])
            psecs.append(s)
            ppairs.append(n)
            same = same_results(plain.results(), ref[k - W])
            identical = identical and same
            log(f"[c5] plain fp64 step {k - W}: {1e3 * s:.1f} ms, identical {same}")
        plain_out = {"value": sum(ppairs) / sum(psecs), "ms_per_step": 1e3 * sum(psecs) / n_plain, "steps": n_plain}
        # the reference path's algorithmic bytes of one step (no gate: every
        # pair traversed), at the last timed state
        plain.prepare(last)
        plain.ctx.set_touch_stats(True)
        plain.ctx.step_run(scenes.DT, True)
        plain.ctx.set_touch_stats(False)
        _, run_counts = plain.ctx.touch_stats()
        g, lv = plain.ctx.run_info()
        raw_hits = int(g[0]) + int(lv[1][0][7])  # discrete contacts before the filter + CCD contacts
        step_bytes = step_alg_bytes(run_counts, pairs_t[-1], raw_hits)
        plain.close()
    if step_bytes is not None:
        ach = step_bytes / (ms_per_step / 1e3) / 1e9
        roofline["step"] = {"alg_bytes": step_bytes, "ms": ms_per_step, "achieved": ach, "frac": ach / HBM_PEAK_GBS,
                            "bytes_model": "distinct (body, node) records of both traversals of the plain path "
                                           "+ 40P per body + 16 B per pair + (11P + 16) B per raw hit, P = 8"}
    cpu = None
    if not args.no_cpu_baseline:
        t0 = time.time()
        cpu = cpu_baseline(sc.trees, sc.tree_of_body, cpu_rec, cpu_pairs, cpu_mt)
        log(f"[c5] cpu baseline in {time.time() - t0:.1f}s: {cpu['value']:.1f} pairs/s on {cpu['cores']} cores, "
            f"{cpu['value_1_thread']:.1f} on 1")
    del sc, states
    c2 = None
    if args.c2 and time.time() - t_start < args.c2_deadline:
        c2 = run_small_figures("c2", args, weights)
    out = {
        "metric": METRIC, "value": value, "unit": "pair queries/s", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64" if prec == d2.D2_FP32 else "f64", "data": "synthetic",
        "form": {"precision": args.precision, "learned_gate": weights is not None,
                 "gate_phases": tuned[3] if weights is not None else None, "graph": False, "weights": wsrc},
        "plain_fp64": plain_out, "speedup_vs_plain_fp64": plain_out["ms_per_step"] / ms_per_step if plain_out else None,
        "identical_to_plain_fp64": identical, "gate": gate, "parts": parts,
        "config": {"workload": f"{args.bodies} subdivision-{args.subdiv} icospheres U(0.5,1.5) ({sc_tris(args)} "
                               f"triangles) in and above a static 40 m funnel (32768 triangles) at 0.1 bodies/m^3, "
                               f"v ~ N(0,1) - 3z m/s, dt = 1/60; per step: device broad phase + clusters over the "
                               f"step's states, then CCD + step rule + cluster steps + discrete + filter; ballistic "
                               f"motion between steps",
                   "bodies": args.bodies + 1, "triangles": sc_tris(args), "pairs_per_step": pairs_t},
        "roofline": roofline, "cpu_baseline": cpu,
        "end_to_end": {"ms_per_step": e2e_ms, "value": pairs_t[-1] / (e2e_ms / 1e3), "unit": "pair queries/s",
                       "includes": "H2D of body states, device broad phase + clusters, the step, D2H of contacts and "
                                   "per-pair CCD values"},
        "device_pipeline": {"broad_phase_ms": float(fast_prep[:, 0].mean()), "clusters_ms": float(fast_prep[:, 1].mean())},
        "setup_s": {"scene": build_s, "tree_upload": fast.upload_s},
        "c2": c2,
    }
    print(json.dumps(out), flush=True)


def sc_tris(args):
    return 32768 + args.bodies * 20 * 4 ** args.subdiv


# --- the 4096-body and 100-body workloads --------------------------------------------
def make_ctx(scene, prec, graph, weights=None, clusters=None, tune=True):
    """A context holding the scene; launch shapes (and the gate's phase mask)
    picked by the library's autotuner on this step."""
    import delta2_amd as d2
    from delta2_amd import scenes
    ctx = d2.Context(0, prec)
    ctx.set_pruning(prec == d2.D2_FP32)  # FP64 contexts are the plain path: every node pair of the reference
    ctx.upload_trees(scene.trees)
    ctx.upload_bodies(scene.tree_of_body, scene.records)
    ctx.upload_pairs(scene.pairs)
    cl, n_cl = clusters
    ctx.upload_clusters(cl, n_cl)
    if weights is not None:
        ctx.mlp_set_params(weights)
        ctx.set_learned(True)
    ctx.set_graph(graph)
    ctx.step_run(scenes.DT, True)  # sizes the buffers, observes the level counts
    if tune:
        shape, ms = ctx.autotune("step", reps=3)
        log(f"[bench] autotuned {shape}: {ms:.3f} ms")
    return ctx


def timed_steps(ctx, dt, steps, warmup):
    for _ in range(warmup):
        ctx.step_run(dt, True)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.step_run(dt, True)
    ctx.sync()
    return (time.perf_counter() - t0) / steps


def evolving_steps(ctx, scene, steps=20):
    """A stepping sequence whose inputs change every step: the states advance
    ballistically, then the device broad phase, clusters and the step run on
    them (the pair list and clusters differ from step to step).  Returns the
    step time and how many graph segments were captured on the way."""
    from delta2_amd import scenes
    rec = scene.records
    c0, l0 = ctx.graph_stats()
    secs, pairs = [], []
    for _ in range(steps):
        rec = scenes.advance_gravity(rec)
        ctx.upload_bodies(scene.tree_of_body, rec)
        ctx.project_future(np.minimum(2.0 * rec[:, 39], scenes.DT))
        pairs.append(ctx.broad_phase(mode=1))
        ctx.build_clusters()
        ctx.sync()
        t0 = time.perf_counter()
        ctx.step_run(scenes.DT, True)
        ctx.sync()
        secs.append(time.perf_counter() - t0)
    c1, l1 = ctx.graph_stats()
    return {"steps": steps, "ms_per_step": 1e3 * float(np.mean(secs)), "captures": c1 - c0, "graph_launches": l1 - l0,
            "pairs": [int(min(pairs)), int(max(pairs))]}


def run_small_figures(workload, args, weights, steps=None, tune=True):
    """One scene's step replayed: the plain FP64 path and this form (graph
    replay, autotuned shapes), their results compared."""
    import delta2_amd as d2
    from delta2_amd import scenes
    t0 = time.time()
    scene = scenes.cached(workload, scenes.scene_c2 if workload == "c2" else scenes.scene_c1)
    log(f"[{workload}] {scene.n_bodies} bodies, {scene.n_triangles()} triangles, {scene.pairs.shape[0]} pairs, "
        f"built in {time.time() - t0:.1f}s")
    steps = steps or min(args.steps, 10)
    clusters = scenes.clusters(scene.pairs, scene.n_bodies, scene.records)
    prec = d2.D2_FP64 if args.precision == "fp64" else d2.D2_FP32
    plain = make_ctx(scene, d2.D2_FP64, True, None, clusters, tune=tune)
    plain_sec = timed_steps(plain, scenes.DT, steps, 2)
    plain_res = step_results(plain)
    plain.close()
    ctx = make_ctx(scene, prec, args.graph == "on", weights, clusters, tune=tune)
    sec = timed_steps(ctx, scenes.DT, steps, 2)
    identical = same_results(step_results(ctx), plain_res)
    gate = ctx.gate_counts() if weights is not None else None
    # the same form on a sequence whose pair list changes every step, with
    # graph replay and eagerly: how often it recaptures, and what that costs
    evolving = {"graph": evolving_steps(ctx, scene)}
    ctx.set_graph(False)
    evolving["eager"] = evolving_steps(ctx, scene)
    ctx.close()
    log(f"[{workload}] evolving 20 steps: graph {evolving['graph']['ms_per_step']:.3f} ms/step with "
        f"{evolving['graph']['captures']} captures, eager {evolving['eager']['ms_per_step']:.3f} ms/step")
    n = int(scene.pairs.shape[0])
    log(f"[{workload}] plain fp64 {1e3 * plain_sec:.3f} ms/step, this form {1e3 * sec:.3f} ms/step, identical {identical}")
    return {"value": n / sec, "ms_per_step": 1e3 * sec, "steps": steps, "pairs": n,
            "plain_fp64": {"value": n / plain_sec, "ms_per_step": 1e3 * plain_sec},
            "speedup_vs_plain_fp64": plain_sec / sec, "identical_to_plain_fp64": identical, "gate": gate,
            "graph": args.graph == "on", "evolving": evolving,
            "workload": "4096 bodies: 2048 subdiv-3 icospheres U(0.5,1.5) + 2048 hulls of 64 ellipsoid points; "
                        "Poisson-disk 30% packing; v~N(0,2); dt=1/60; one step replayed"
            if workload == "c2" else "100 subdiv-2 icospheres U(0.5,1.5) in a 10 m box; v~N(0,1); dt=1/60"}


def run_small(args):
    """--workload c1 / c2: the one-step figures plus roofline and CPU baseline."""
    import delta2_amd as d2
    from delta2_amd import scenes
    weights, wsrc = gate_weights() if args.learned == "on" else (None, None)
    fig = run_small_figures(args.workload, args, weights, steps=args.steps, tune=not args.no_autotune)
    scene = scenes.cached(args.workload, scenes.scene_c2 if args.workload == "c2" else scenes.scene_c1)
    prec = d2.D2_FP64 if args.precision == "fp64" else d2.D2_FP32
    clusters = scenes.clusters(scene.pairs, scene.n_bodies, scene.records)
    ctx = make_ctx(scene, prec, False, weights, clusters, tune=not args.no_autotune)
    ctx.reset_stats()
    ctx.set_profiling(True)
    for _ in range(args.profile_steps):
        ctx.step_run(scenes.DT, True)
    ctx.sync()
    ctx.set_profiling(False)
    stats = kernel_table(ctx, args.workload)
    ctx.set_touch_stats(True)
    ctx.step_run(scenes.DT, True)
    ctx.set_touch_stats(False)
    per_level, _ = ctx.touch_stats()
    per_level = per_level * args.profile_steps  # the profiled launches ran that many steps
    roofline = roofline_of(stats, per_level, f"{args.workload}/{args.precision}/"
                                             f"{'learned' if weights is not None else 'plain'}")
    rec = ctx.download_bodies()
    ctx.close()
    cpu = None
    if not args.no_cpu_baseline:
        mt = np.minimum(scene.records[scene.pairs[:, 0], 39], scene.records[scene.pairs[:, 1], 39])
        cpu = cpu_baseline(scene.trees, scene.tree_of_body, rec, scene.pairs, mt)
    out = {"metric": METRIC, "value": fig["value"], "unit": "pair queries/s", "n_gpus": 1, "steps": fig["steps"],
           "warmup": 2, "ms_per_step": fig["ms_per_step"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32+f64" if prec == d2.D2_FP32 else "f64", "data": "synthetic",
           "form": {"precision": args.precision, "learned_gate": weights is not None, "graph": args.graph == "on",
                    "weights": wsrc},
           "plain_fp64": fig["plain_fp64"], "speedup_vs_plain_fp64": fig["speedup_vs_plain_fp64"],
           "identical_to_plain_fp64": fig["identical_to_plain_fp64"], "gate": fig["gate"], "evolving": fig["evolving"],
           "config": {"workload": fig["workload"], "bodies": scene.n_bodies, "triangles": scene.n_triangles(),
                      "pairs": fig["pairs"]},
           "roofline": roofline, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


# --- the held-out learned evaluation and the batched scenes --------------------------
def run_c4(args):
    """Learned vs plain at equal TOI error on the held-out 4096-body C1-density
    scene: identical step results (contacts, per-pair TOI/scale, cluster
    steps) are checked, so the speedup is at zero TOI difference."""
    import delta2_amd as d2
    from delta2_amd import scenes
    t0 = time.time()
    sc = scenes.scene_c4()
    cl = scenes.clusters(sc.pairs, sc.n_bodies, sc.records)
    n_pairs = int(sc.pairs.shape[0])
    log(f"[c4] {sc.n_bodies} bodies, {n_pairs} pairs, built in {time.time() - t0:.1f}s")
    w, src = gate_weights()
    out, ref = {}, None
    for name, prec, wt in (("plain_fp64", d2.D2_FP64, None), ("learned_fp64", d2.D2_FP64, w),
                           ("plain_fp32", d2.D2_FP32, None), ("learned_fp32", d2.D2_FP32, w)):
        ctx = make_ctx(sc, prec, True, wt, cl)
        sec = timed_steps(ctx, scenes.DT, args.steps, args.warmup)
        res = step_results(ctx)
        ref = res if ref is None else ref
        gate = ctx.gate_counts() if wt is not None else None
        out[name] = {"ms_per_step": 1e3 * sec, "value": n_pairs / sec, "identical_to_plain_fp64": same_results(res, ref),
                     "gate": gate}
        log(f"[c4] {name}: {1e3 * sec:.3f} ms/step, identical {out[name]['identical_to_plain_fp64']}, gate {gate}")
        ctx.close()
    best = out["learned_fp32"]
    res = {"metric": METRIC, "value": best["value"], "unit": "pair queries/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": best["ms_per_step"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
           "config": {"workload": "held-out 4096 subdiv-2 icospheres U(0.5,1.5) at the C1 density (34.5 m box), "
                                  "v~N(0,1), dt=1/60; one step (CCD+step rule+cluster steps+discrete+filter)",
                      "pairs": n_pairs, "weights": src},
           "variants": out,
           "learned_vs_plain_fp64": best["value"] / out["plain_fp64"]["value"],
           "learned_vs_plain_same_precision": {"fp64": out["learned_fp64"]["value"] / out["plain_fp64"]["value"],
                                               "fp32": best["value"] / out["plain_fp32"]["value"]}}
    print(json.dumps(res), flush=True)


def run_c3(args):
    """Batched workload: 256 independent scenes (16-512 bodies, C1 meshes).
    Variants on the same inputs, each a full step over every scene:
      batched_fp32_graph : one d2_step_run over all scenes' pairs, FP32 cull,
                           graph replay, autotuned launch shapes (the value)
      batched_fp64_graph : the same in FP64, default shapes
      plain_fp64_scenes  : the plain path as the reference calls it, one
                           scene at a time (pairs + clusters uploaded per
                           scene, eager FP64 d2_step_run)"""
    import delta2_amd as d2
    from delta2_amd import scenes
    t0 = time.time()
    sc = scenes.scene_c3(n_scenes=args.scenes)
    n_pairs = int(sc.pairs.shape[0])
    log(f"[c3] {len(sc.counts)} scenes, {sc.n_bodies} bodies, {sc.n_triangles()} triangles, {n_pairs} pairs, "
        f"built in {time.time() - t0:.1f}s")
    cl, n_cl = scenes.clusters(sc.pairs, sc.n_bodies, sc.records)
    out = {}
    for name, prec, graph in (("batched_fp32_graph", d2.D2_FP32, True), ("batched_fp64_graph", d2.D2_FP64, True)):
        ctx = d2.Context(0, prec)  # capacities grow on demand
        ctx.set_pruning(prec == d2.D2_FP32)
        ctx.upload_trees(sc.trees)
        ctx.upload_bodies(sc.tree_of_body, sc.records)
        ctx.upload_pairs(sc.pairs)
        ctx.upload_clusters(cl, n_cl)
        ctx.set_graph(graph)
        ctx.step_run(scenes.DT, True)  # observe the level counts
        tune = None
        if prec == d2.D2_FP32:
            tune, ms = ctx.autotune("step", reps=3)
            log(f"[c3] autotuned shape {tune}: {ms:.3f} ms")
        sec = timed_steps(ctx, scenes.DT, args.steps, args.warmup)
        out[name] = {"ms_per_step": 1e3 * sec, "value": n_pairs / sec, "tune": tune or ctx.get_tune()}
        log(f"[c3] {name}: {1e3 * sec:.3f} ms/step, {n_pairs / sec:.0f} pair queries/s")
        if prec == d2.D2_FP32:
            ctx.set_graph(False)
            ctx.reset_stats()
            ctx.set_profiling(True)
            for _ in range(args.profile_steps):
                ctx.step_run(scenes.DT, True)
            ctx.sync()
            ctx.set_profiling(False)
            kernel_table(ctx, "c3")
        ctx.close()
    # plain: one scene at a time
    ctx = d2.Context(0, d2.D2_FP64)
    ctx.set_pruning(False)
    ctx.upload_trees(sc.trees)
    ctx.upload_bodies(sc.tree_of_body, sc.records)
    sob = sc.scene_of_body[sc.pairs[:, 0]]
    per_scene = []
    for k in range(len(sc.counts)):
        idx = np.nonzero(sob == k)[0]
        if idx.size:
            used, lcl = np.unique(cl[idx], return_inverse=True)
            per_scene.append((sc.pairs[idx], lcl.astype(np.int32), used.size))
    reps = max(1, args.steps // 4)

    def plain_pass():
        for p, c_, n in per_scene:
            ctx.upload_pairs(p)
            ctx.upload_clusters(c_, n)
            ctx.step_run(scenes.DT, True)
    plain_pass()
    ctx.sync()
    t1 = time.perf_counter()
    for _ in range(reps):
        plain_pass()
    ctx.sync()
    sec = (time.perf_counter() - t1) / reps
    ctx.close()
    out["plain_fp64_scenes"] = {"ms_per_step": 1e3 * sec, "value": n_pairs / sec}
    log(f"[c3] plain_fp64_scenes: {1e3 * sec:.3f} ms/step, {n_pairs / sec:.0f} pair queries/s")
    best = out["batched_fp32_graph"]
    res = {
        "metric": METRIC, "value": best["value"], "unit": "pair queries/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": best["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": f"{len(sc.counts)} independent scenes, log-uniform 16-512 bodies each, subdiv-2 "
                               "icospheres U(0.5,1.5), C1 density, v~N(0,1), dt=1/60; one step over all scenes",
                   "bodies": sc.n_bodies, "triangles": sc.n_triangles(), "pairs": n_pairs, "clusters": n_cl},
        "variants": out,
        "speedup_vs_plain_fp64_scenes": best["value"] / out["plain_fp64_scenes"]["value"],
        "speedup_vs_batched_fp64": best["value"] / out["batched_fp64_graph"]["value"],
    }
    print(json.dumps(res), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c5", choices=["c1", "c2", "c3", "c4", "c5"])
    ap.add_argument("--bodies", type=int, default=65536, help="c5: bodies besides the funnel")
    ap.add_argument("--subdiv", type=int, default=3, help="c5: icosphere subdivision (3: 1280 triangles per body)")
    ap.add_argument("--plain-steps", type=int, default=3, help="c5: timed steps of the plain FP64 path (0: all)")
    ap.add_argument("--c2", type=int, default=1, help="c5: also measure the 4096-body workload (extra key)")
    ap.add_argument("--c2-deadline", type=float, default=330.0, help="c5: skip the 4096-body figures after this many s")
    ap.add_argument("--tune-budget", type=float, default=45.0, help="c5: seconds spent tuning launch shapes")
    ap.add_argument("--scenes", type=int, default=256, help="c3: number of independent scenes")
    ap.add_argument("--precision", default="fp32", choices=["fp64", "fp32"])
    ap.add_argument("--learned", default="on", choices=["on", "off"],
                    help="the learned pair gate (weights from delta2_amd/weights; trained here when absent)")
    ap.add_argument("--graph", default="on", choices=["on", "off"], help="c1-c4: graph replay of the step")
    ap.add_argument("--profile-steps", type=int, default=5, help="eager steps timed per kernel for the roofline")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="default launch shapes (profiling runs)")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the plain FP64 reference leg (profiling runs of one configuration)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (HBM counter calibration for rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the path has no multi-GPU sharding",
              file=sys.stderr)
        sys.exit(2)
    import delta2_amd as d2
    d2.lib()
    {"c3": run_c3, "c4": run_c4, "c5": run_pile}.get(args.workload, run_small)(args)


if __name__ == "__main__":
    main()
