"""Benchmark: narrow+CCD pair queries/s on one MI355X.

A step is one pass of the contact-detection path over the workload's
broad-phase pair list with inputs resident in HBM (d2_step_run):
  CCD (compareTreesFullContinuous over [0, min(last steps)]) + the TOI -> step
  rule, the clamped cluster steps, the future states at those steps, then the
  discrete narrow phase (compareTreesFull) + filterContacts,
for every pair.  value = pairs * steps / wall time of the timed steps.  The
timed steps replay one captured HIP graph per step (--graph off: eager
launches); a separate eager pass with HIP events brackets every kernel for the
roofline figures.

The workload is the 4096-body scene (2048 subdivision-3 icospheres scaled
U(0.5,1.5) and 2048 convex hulls of 64 ellipsoid points, Poisson-disk placed
at 30% packing of the bounding spheres, velocities N(0,2), dt = 1/60 s).

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

# Algorithmic HBM bytes per unit of work of each kernel (unit = the item the
# kernel's launch counts): what one unit must read or write once, per-pair
# records (poses, shared by all of a pair's units) not counted.
#   tri_mask   : one node pair: its record (24 B), two node records (64 B),
#                the pair's body/tree ids (16 B), the triangles (2 x 576 B)
#                and eps (2 x 64 B) of both buckets; hit flag and expansion
#                descriptor out (8 B); its kept items out (counted under tri_dist)
#   tri_enum   : no launch of its own (the enumeration is part of tri_mask);
#                its item count is the discrete node pairs received
#   tri_dist   : one triangle pair: item (20 B: node pair, pair, buckets),
#                two triangles (144 B), two eps (16 B)
#   soup_cull  : one soup node pair: as tri_mask (swept boxes of both buckets)
#   soup_point / soup_edge : one sub-test: item (8 B), two triangles (144 B), one 64-B result
#   leaf_point / leaf_edge : one sub-test: item (8 B), two triangles (144 B)
#   soup_reduce: one triangle pair: 15 results (960 B)
ALG_BYTES = {"tri_mask": 1392.0, "tri_enum": 12.0, "tri_dist": 180.0, "soup_cull": 1400.0, "soup_point": 216.0,
             "soup_edge": 216.0, "leaf_point": 152.0, "leaf_edge": 152.0, "soup_reduce": 960.0}
# Estimated FP32 operations per unit of the FP32 cull (tri_mask): per node
# pair 16 triangle preparations (~80 ops) and 64 pair classifications (~130 ops).
ALU_OPS = {"tri_mask": 16 * 80 + 64 * 130}
FP32_PEAK_TFLOPS = 157.3
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(scene, max_time, budget_s=12.0):
    """The oracle (C++ restatement of the reference path) on the host cores,
    on a bounded random sample of the same pairs."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import d2o
    d2o.lib()
    osc = d2o.Scene(scene.trees, scene.tree_of_body, scene.records)
    pairs = scene.pairs
    rng = np.random.default_rng(0)
    order = rng.permutation(pairs.shape[0])
    # calibrate on a small sample, then size the sample to the budget
    n = min(16, pairs.shape[0])
    t1, _ = osc.time_pass(pairs[order[:n]], max_time[order[:n]], threads=1, repeats=1)
    per_pair = max(t1 / n, 1e-6)
    n = int(min(pairs.shape[0], max(n, budget_s / 3.0 / per_pair)))
    sel = order[:n]
    t1, _ = osc.time_pass(pairs[sel], max_time[sel], threads=1, repeats=3)
    cores = len(os.sched_getaffinity(0))
    tn, _ = osc.time_pass(pairs[sel], max_time[sel], threads=cores, repeats=3)
    return {"value": n / tn, "unit": "pair queries/s", "cores": cores, "kind": "port",
            "value_1_thread": n / t1,
            "sample": f"{n} of {pairs.shape[0]} pairs drawn uniformly (seed 0), CCD+step rule+discrete+filter, "
                      f"median of 3 passes; oracle built with g++ -O2 -ffp-contract=off"}


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


def make_ctx(scene, prec, graph, weights=None, clusters=None, tune=True):
    """A context holding the scene; launch shapes (and the gate's phase mask)
    picked by the library's autotuner on this step."""
    import delta2_amd as d2
    from delta2_amd import scenes
    ctx = d2.Context(0, prec)
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


def run_c4(args):
    """Learned vs plain at equal TOI error on the held-out 4096-body C1-density
    scene: identical step results (contacts, per-pair TOI/scale, cluster
    steps) are checked, so the speedup is at zero TOI difference."""
    import delta2_amd as d2
    from delta2_amd import learned, scenes
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


def run_c5(args):
    """Large dense pile: 65536 C1-mesh bodies at the C1 density in and above a
    static 40 m funnel, falling.  Each step: ballistic advance under gravity
    (no contact response: the solver is outside this path), broad phase and
    clusters on the host, then the timed d2_step_run (inputs resident).  The
    plain FP64 path and the learned FP32 form run the same steps; their step
    results are compared every step."""
    import delta2_amd as d2
    from delta2_amd import scenes
    t0 = time.time()
    sc = scenes.scene_c5()
    log(f"[c5] {sc.n_bodies} bodies ({sc.n_triangles()} triangles), built in {time.time() - t0:.1f}s")
    w, src = gate_weights()
    steps = []
    rec = sc.records
    for k in range(args.steps):
        pairs = scenes.c5_pairs(rec, sc.radii) if k else sc.pairs
        steps.append((rec, pairs, scenes.clusters(pairs, sc.n_bodies, rec)))
        rec = scenes.advance_gravity(rec)
    out, ref = {}, {}
    for name, prec, wt in (("plain_fp64", d2.D2_FP64, None), ("learned_fp32", d2.D2_FP32, w)):
        ctx = d2.Context(0, prec)
        ctx.upload_trees(sc.trees)
        if wt is not None:
            ctx.mlp_set_params(wt)
            ctx.set_learned(True)
        total, n_total, same = 0.0, 0, True
        cl_steps = []
        for k, (r, pairs, cl) in enumerate(steps):
            ctx.upload_bodies(sc.tree_of_body, r)
            ctx.upload_pairs(pairs)
            ctx.upload_clusters(*cl)
            if k == 0:
                ctx.step_run(scenes.DT, True)  # sizes buffers and capacities
                if wt is not None:
                    shape, ms = ctx.autotune("step", reps=1)
                    log(f"[c5] autotuned {shape}: {ms:.1f} ms")
                ctx.upload_bodies(sc.tree_of_body, r)
            ctx.sync()
            t1 = time.perf_counter()
            _, st = ctx.step_run(scenes.DT, True, want_steps=True)
            ctx.sync()
            total += time.perf_counter() - t1
            n_total += pairs.shape[0]
            res = step_results(ctx)
            if name == "plain_fp64":
                ref[k] = (res, st)
            else:
                same = same and same_results(res, ref[k][0]) and np.array_equal(st, ref[k][1])
            cl_steps.append(st)
        allst = np.concatenate(cl_steps)
        out[name] = {"ms_per_step": 1e3 * total / len(steps), "value": n_total / total,
                     "identical_to_plain_fp64": same, "gate": ctx.gate_counts() if wt is not None else None,
                     "cluster_step_range_s": [float(allst.min()), float(allst.max())]}
        log(f"[c5] {name}: {out[name]['ms_per_step']:.1f} ms/step, identical {same}")
        ctx.close()
    # the same steps with broad phase and clusters on the device, as the
    # reference's cluster stepper forms them (broadPhaseEmbree(particles): the
    # static funnel pairs with every body inside its box)
    ctx = d2.Context(0, d2.D2_FP32)
    ctx.upload_trees(sc.trees)
    ctx.mlp_set_params(w)
    ctx.set_learned(True)
    dev = {"broad_phase_ms": [], "clusters_ms": [], "step_ms": [], "pairs": [], "clusters": []}
    for k, (r, _, _) in enumerate(steps):
        for rep in range(2 if k == 0 else 1):  # the first pass sizes buffers
            ctx.upload_bodies(sc.tree_of_body, r)
            ctx.project_future(np.minimum(2.0 * r[:, 39], scenes.DT))  # broad_phase_embree_cluster.cpp:121-127
            ctx.sync()
            t1 = time.perf_counter()
            n_pairs = ctx.broad_phase(mode=1)
            t2 = time.perf_counter()
            n_cl = ctx.build_clusters()
            t3 = time.perf_counter()
            ctx.step_run(scenes.DT, True)
            ctx.sync()
            t4 = time.perf_counter()
        dev["broad_phase_ms"].append(1e3 * (t2 - t1))
        dev["clusters_ms"].append(1e3 * (t3 - t2))
        dev["step_ms"].append(1e3 * (t4 - t3))
        dev["pairs"].append(int(n_pairs))
        dev["clusters"].append(int(n_cl))
    ctx.close()
    dev["pair_queries_per_s"] = sum(dev["pairs"]) / (sum(dev["broad_phase_ms"]) + sum(dev["clusters_ms"]) +
                                                     sum(dev["step_ms"])) * 1e3
    log(f"[c5] device pipeline: pairs {dev['pairs']}, broad phase {dev['broad_phase_ms']} ms, clusters "
        f"{dev['clusters_ms']} ms, step {dev['step_ms']} ms")
    out["device_pipeline_fp32"] = dev
    best = out["learned_fp32"]
    res = {"metric": METRIC, "value": best["value"], "unit": "pair queries/s", "n_gpus": 1, "steps": args.steps,
           "warmup": 1, "ms_per_step": best["ms_per_step"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
           "config": {"workload": "65536 subdiv-2 icospheres U(0.5,1.5) (21 M triangles) at 0.1 bodies/m^3 in and above "
                                  "a static 40 m funnel (32768 triangles), v ~ N(0,1) - 3 z m/s, dt = 1/60, "
                                  f"{args.steps} ballistic steps; broad phase on the host (untimed)",
                      "pairs_per_step": [int(p.shape[0]) for _, p, _ in steps], "weights": src},
           "variants": out, "learned_fp32_vs_plain_fp64": best["value"] / out["plain_fp64"]["value"]}
    print(json.dumps(res), flush=True)


def timed_steps(ctx, dt, steps, warmup):
    for _ in range(warmup):
        ctx.step_run(dt, True)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.step_run(dt, True)
    ctx.sync()
    return (time.perf_counter() - t0) / steps


def end_to_end(ctx, scene, clusters, steps=5):
    """The same step with host buffers on both sides: bodies, pairs and
    clusters uploaded, contacts and per-pair CCD values downloaded."""
    from delta2_amd import scenes
    cl, n_cl = clusters
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.upload_bodies(scene.tree_of_body, scene.records)
        ctx.upload_pairs(scene.pairs)
        ctx.upload_clusters(cl, n_cl)
        ctx.step_run(scenes.DT, True)
        ctx.result(with_pairs=True)
    sec = (time.perf_counter() - t0) / steps
    return {"value": scene.pairs.shape[0] / sec, "unit": "pair queries/s", "ms_per_step": 1e3 * sec, "steps": steps,
            "includes": "H2D of body records, pairs, clusters; D2H of contacts and per-pair CCD values"}


def device_pipeline(scene, prec, weights, steps=5):
    """Broad phase, cluster separation and the step, all on the device, over
    the scene's bodies (the pair list never leaves the GPU)."""
    import delta2_amd as d2
    from delta2_amd import scenes
    ctx = d2.Context(0, prec)
    ctx.upload_trees(scene.trees)
    ctx.upload_bodies(scene.tree_of_body, scene.records)
    ctx.project_future(np.minimum(2.0 * scene.records[:, 39], scenes.DT))  # broad_phase_embree_cluster.cpp:121-127
    if weights is not None:
        ctx.mlp_set_params(weights)
        ctx.set_learned(True)
    ctx.set_graph(True)
    t_bp = t_cl = t_st = 0.0
    for k in range(steps + 1):
        t0 = time.perf_counter()
        n_pairs = ctx.broad_phase(mode=1)
        t1 = time.perf_counter()
        n_cl = ctx.build_clusters()
        t2 = time.perf_counter()
        ctx.step_run(scenes.DT, True)
        ctx.sync()
        t3 = time.perf_counter()
        if k:  # the first step sizes buffers and captures graphs
            t_bp += t1 - t0
            t_cl += t2 - t1
            t_st += t3 - t2
    ctx.close()
    ms = lambda x: 1e3 * x / steps  # noqa: E731
    total = ms(t_bp + t_cl + t_st)
    return {"bodies": scene.n_bodies, "pairs": n_pairs, "clusters": n_cl, "broad_phase_ms": ms(t_bp),
            "clusters_ms": ms(t_cl), "step_ms": ms(t_st), "total_ms": total,
            "bodies_per_s_broad_phase": scene.n_bodies / (t_bp / steps), "pair_queries_per_s": n_pairs / (total / 1e3)}


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
            for k, (n, ms, items) in sorted(ctx.kernel_stats().items(), key=lambda kv: -kv[1][1]):
                if n:
                    log(f"[c3]   {k:12s} launches {n:6d}  total {ms:9.3f} ms  avg {ms / n * 1e3:9.2f} us  items {items:.0f}")
        ctx.close()
    # plain: one scene at a time
    ctx = d2.Context(0, d2.D2_FP64)
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
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c2", choices=["c1", "c2", "c3", "c4", "c5"])
    ap.add_argument("--scenes", type=int, default=256, help="c3: number of independent scenes")
    ap.add_argument("--precision", default="fp32", choices=["fp64", "fp32"])
    ap.add_argument("--learned", default="on", choices=["on", "off"],
                    help="the learned pair gate (weights from delta2_amd/weights; trained here when absent)")
    ap.add_argument("--graph", default="on", choices=["on", "off"])
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
    from delta2_amd import scenes
    d2.lib()
    if args.workload == "c3":
        run_c3(args)
        return
    if args.workload == "c4":
        run_c4(args)
        return
    if args.workload == "c5":
        run_c5(args)
        return
    t0 = time.time()
    scene = scenes.cached(args.workload, scenes.scene_c2 if args.workload == "c2" else scenes.scene_c1)
    log(f"[bench] scene {scene.name}: {scene.n_bodies} bodies, {scene.n_triangles()} triangles, "
        f"{scene.pairs.shape[0]} pairs, built in {time.time() - t0:.1f}s")
    pairs = scene.pairs
    max_time = np.minimum(scene.records[pairs[:, 0], 39], scene.records[pairs[:, 1], 39])
    prec = d2.D2_FP64 if args.precision == "fp64" else d2.D2_FP32
    clusters = scenes.clusters(pairs, scene.n_bodies, scene.records)
    n_cl = clusters[1]
    weights, wsrc = gate_weights() if args.learned == "on" else (None, None)
    # the plain FP64 path on the same inputs (reference point of the speedup)
    plain_sec, plain_res = None, None
    if not args.no_plain:
        plain = make_ctx(scene, d2.D2_FP64, True, None, clusters, tune=not args.no_autotune)
        plain_sec = timed_steps(plain, scenes.DT, args.steps, args.warmup)
        plain_res = step_results(plain)
        plain.close()
    ctx = make_ctx(scene, prec, args.graph == "on", weights, clusters, tune=not args.no_autotune)
    n_contacts = ctx.step_run(scenes.DT, True)
    sec = timed_steps(ctx, scenes.DT, args.steps, args.warmup)
    identical = same_results(step_results(ctx), plain_res) if plain_res else None
    ms_per_step = 1e3 * sec
    value = pairs.shape[0] / sec
    gate = ctx.gate_counts() if weights is not None else None
    log(f"[bench] plain fp64: {1e3 * plain_sec if plain_sec else float('nan'):.3f} ms/step; this form: {ms_per_step:.3f} ms/step "
        f"({n_cl} clusters, {n_contacts} filtered contacts, graph {args.graph}, gate {gate}, identical {identical})")
    # per-kernel times: eager steps, every kernel bracketed by HIP events on the context's stream
    ctx.set_graph(False)
    ctx.reset_stats()
    ctx.set_profiling(True)
    for _ in range(args.profile_steps):
        ctx.step_run(scenes.DT, True)
    ctx.sync()
    ctx.set_profiling(False)
    stats = ctx.kernel_stats()
    for k, (n, ms, items) in sorted(stats.items(), key=lambda kv: -kv[1][1]):
        if n:
            log(f"[bench]   {k:12s} launches {n:6d}  total {ms:9.3f} ms  avg {ms / n * 1e3:9.2f} us  items {items:.0f}")

    dom = max((k for k in stats if k in ALG_BYTES and stats[k][0]), key=lambda k: stats[k][1])
    n, ms, items = stats[dom]
    avg_s = ms / n / 1e3
    bytes_per_launch = ALG_BYTES[dom] * items / n
    achieved = bytes_per_launch / avg_s / 1e9
    traffic = None
    if os.path.exists(TRAFFIC_FILE):  # measured HBM bytes per dispatch (rocprofv3 FETCH_SIZE + WRITE_SIZE, calibrated)
        with open(TRAFFIC_FILE) as f:
            key = f"{scene.name}/{args.precision}/{'learned' if weights is not None else 'plain'}"
            t = json.load(f).get(key, {}).get(dom)
        traffic = t.get("bytes_per_dispatch") if t else None
    roofline = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                "avg_launch_us": avg_s * 1e6, "alg_bytes_per_launch": bytes_per_launch}
    if dom in ALU_OPS:
        tf = ALU_OPS[dom] * items / n / avg_s / 1e12
        roofline["alu"] = {"achieved_tflops_est": tf, "peak_fp32_tflops": FP32_PEAK_TFLOPS, "frac": tf / FP32_PEAK_TFLOPS,
                           "ops_per_unit_est": ALU_OPS[dom]}
    ctx.set_graph(args.graph == "on")
    e2e = end_to_end(ctx, scene, clusters)
    log(f"[bench] end to end (host buffers): {e2e['ms_per_step']:.3f} ms/step")
    dev = device_pipeline(scene, prec, weights)
    log(f"[bench] device pipeline: broad phase {dev['broad_phase_ms']:.3f} ms ({dev['pairs']} pairs), clusters "
        f"{dev['clusters_ms']:.3f} ms ({dev['clusters']}), step {dev['step_ms']:.3f} ms")
    cpu = None
    if not args.no_cpu_baseline:
        t0 = time.time()
        cpu = cpu_baseline(scene, max_time)
        log(f"[bench] cpu baseline in {time.time() - t0:.1f}s: {cpu['value']:.1f} pairs/s on {cpu['cores']} cores, "
            f"{cpu['value_1_thread']:.1f} on 1")
    out = {
        "metric": METRIC, "value": value, "unit": "pair queries/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if prec == d2.D2_FP64 else "f32+f64", "data": "synthetic",
        "form": {"precision": args.precision, "learned_gate": args.learned == "on", "graph": args.graph == "on",
                 "weights": wsrc},
        "plain_fp64": {"value": pairs.shape[0] / plain_sec, "ms_per_step": 1e3 * plain_sec} if plain_sec else None,
        "speedup_vs_plain_fp64": plain_sec / sec if plain_sec else None, "identical_to_plain_fp64": identical,
        "gate": gate,
        "config": {"workload": "4096 bodies: 2048 subdiv-3 icospheres U(0.5,1.5) + 2048 hulls of 64 ellipsoid points; "
                               "Poisson-disk 30% packing; v~N(0,2); dt=1/60; CCD+step rule then discrete+filter"
                   if scene.name == "c2" else "100 subdiv-2 icospheres U(0.5,1.5) in a 10 m box; v~N(0,1); dt=1/60",
                   "bodies": scene.n_bodies, "triangles": scene.n_triangles(), "pairs": int(pairs.shape[0]),
                   "precision": args.precision},
        "roofline": roofline, "cpu_baseline": cpu, "end_to_end": e2e, "device_pipeline": dev,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
