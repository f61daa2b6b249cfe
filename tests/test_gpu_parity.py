"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

FP64 is required to reproduce the oracle: identical contact index tuples in
identical order, and reals
  * bit for bit where the device works on the same inputs as the oracle (the
    discrete phase over uploaded future states): the kernels evaluate the
    reference's arithmetic in the oracle's order without contraction;
  * within 1e-9 relative (|x - y| <= 1e-9 |y|) where the device first computes
    poses itself (the CCD projects the future states with the device's sin
    and cos, which may differ from glibc's in the last bit).
Per-pair step scales and min_time_out are held to the same relative bound.
"""
import numpy as np
import pytest

import d2o
import delta2_amd as d2
from gpu_cases import c1_scene, dense_scene, hull_scene, oracle_scene, sparse_scene, with_future
from delta2_amd import scenes

pytestmark = pytest.mark.gpu
TOL = 1e-9


def close(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b) <= TOL * np.abs(b)


def assert_same(gpu, ora_off, ora_ints, ora_reals, label, exact=True):
    assert np.array_equal(gpu.offsets, ora_off), f"{label}: per-pair contact counts differ"
    assert np.array_equal(gpu.ints, ora_ints), f"{label}: contact index tuples differ"
    if exact:
        bad = gpu.reals != ora_reals
        assert not bad.any(), f"{label}: {np.count_nonzero(bad)} reals differ in their bits " \
                              f"(max abs diff {np.abs(gpu.reals - ora_reals).max()})"
        return
    ok = close(gpu.reals, ora_reals)
    same = np.count_nonzero(gpu.reals == ora_reals)
    print(f"[{label}] {same} of {ora_reals.size} reals bit-identical", flush=True)
    assert ok.all(), f"{label}: {np.count_nonzero(~ok)} reals beyond 1e-9 relative " \
                     f"(max abs diff {np.abs(gpu.reals - ora_reals).max()})"


def _fingerprint():
    out = {}
    for name, make in SCENES:
        if make.cache_info().currsize:
            sc = make()
            out[name] = [(t[0].shape, t[1].shape, int(t[0].sum()), float(t[1].sum())) for t in sc.trees]
    return out


@pytest.fixture(autouse=True)
def scenes_untouched(request):
    """Cached scene arrays must be bit-stable across every test."""
    before = _fingerprint()
    yield
    after = _fingerprint()
    for k, v in before.items():
        assert after[k] == v, f"{request.node.name} modified the cached {k} scene"


@pytest.fixture(scope="module")
def ctx():
    c = d2.Context(0, d2.D2_FP64)
    yield c
    c.close()


def load(ctx, s, records):
    ctx.upload_trees(s.trees)
    ctx.upload_bodies(s.tree_of_body, records)


# c2_256: the 4096-body mixed workload's generator at 256 bodies
SCENES = [("c1", c1_scene), ("dense", dense_scene), ("hulls", hull_scene), ("c2_256", sparse_scene)]


@pytest.mark.parametrize("name,make", SCENES)
@pytest.mark.parametrize("filt", [False, True])
def test_narrow_parity(ctx, name, make, filt):
    s = make()
    print(f"[narrow] scene={name} pairs={s.pairs.shape[0]} filter={filt}", flush=True)
    rec = with_future(s.records, scenes.DT)
    load(ctx, s, rec)
    gpu = d2.compare_trees_full(ctx, s.pairs, filter=filt)
    o_off, o_ints, o_reals = oracle_scene(s, rec).narrow(s.pairs, filter=filt)
    assert_same(gpu, o_off, o_ints, o_reals, f"narrow {name} filter={filt}")
    if name == "dense":
        assert gpu.n > 100  # the dense scene must exercise many contacts


@pytest.mark.parametrize("prec", [d2.D2_FP64, d2.D2_FP32])
def test_narrow_repeat_runs(prec):
    """A fresh context's first run launches every level with the default
    grids; later runs size each level's grids (and the post-processing's)
    from the previous run's sizes.  Every run equals the oracle."""
    s = dense_scene()
    rec = with_future(s.records, scenes.DT)
    o_off, o_ints, o_reals = oracle_scene(s, rec).narrow(s.pairs, filter=True)
    c = d2.Context(0, prec)
    try:
        load(c, s, rec)
        for run in range(3):
            print(f"[narrow repeat] precision={prec} run={run}", flush=True)
            gpu = d2.compare_trees_full(c, s.pairs, filter=True)
            assert_same(gpu, o_off, o_ints, o_reals, f"narrow repeat {prec} run {run}")
    finally:
        c.close()


def _stacked_scene():
    """Two pairs of bodies sharing their centres (subdivision-3 and -2
    icospheres): thousands of contacts per pair, beyond the post-processing's
    256-thread block (512) and its one-wavefront LDS pass (2048)."""
    from gpu_cases import _seed
    rng = np.random.default_rng(5)
    meshes = [scenes.icosphere(3, 1.0), scenes.icosphere(3, 1.0), scenes.icosphere(2, 1.0),
              scenes.icosphere(2, 1.02)]
    pos = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 0.0], [5.0, 0.0, 0.0], [5.0, 0.0, 0.0]])
    trees, rec, radii = scenes._bodies(meshes, rng, pos, 0.5, _seed())
    rec[:, 14:28] = rec[:, 0:14]
    s = scenes.Scene("stacked", meshes, trees, rec, radii)
    s.pairs = np.array([[0, 1], [2, 3]], dtype=np.int32)
    return s


@pytest.mark.parametrize("prec", [d2.D2_FP64, d2.D2_FP32])
def test_pairs_with_thousands_of_contacts(prec):
    s = _stacked_scene()
    rec = with_future(s.records, scenes.DT)
    c = d2.Context(0, prec)
    try:
        load(c, s, rec)
        for filt in (False, True):
            print(f"[stacked] precision={prec} narrow filter={filt}", flush=True)
            gpu = d2.compare_trees_full(c, s.pairs, filter=filt)
            o_off, o_ints, o_reals = oracle_scene(s, rec).narrow(s.pairs, filter=filt)
            if not filt:  # one pair per path: over 2048 contacts, and 512-2048
                n = np.diff(o_off)
                assert n[0] > 2048 and 512 < n[1] <= 2048, n
            assert_same(gpu, o_off, o_ints, o_reals, f"stacked narrow {prec} filter={filt}")
        load(c, s, s.records)
        mt = np.full(2, scenes.DT)
        print(f"[stacked] precision={prec} ccd", flush=True)
        gpu = d2.compare_trees_full_continuous(c, s.pairs, mt)
        o_off, o_ints, o_reals, o_pp = oracle_scene(s, s.records).ccd(s.pairs, mt)
        assert np.all(np.diff(o_off) > 2048)
        assert_same(gpu, o_off, o_ints, o_reals, f"stacked ccd {prec}", exact=False)
        assert close(gpu.per_pair, o_pp).all()
    finally:
        c.close()


@pytest.mark.parametrize("name,make", SCENES)
def test_ccd_parity(ctx, name, make):
    s = make()
    print(f"[ccd] scene={name} pairs={s.pairs.shape[0]}", flush=True)
    rec = s.records
    load(ctx, s, rec)
    mt = np.minimum(rec[s.pairs[:, 0], 39], rec[s.pairs[:, 1], 39]) if s.pairs.size else np.zeros(0)
    gpu = d2.compare_trees_full_continuous(ctx, s.pairs, mt)
    o_off, o_ints, o_reals, o_pp = oracle_scene(s, rec).ccd(s.pairs, mt)
    assert_same(gpu, o_off, o_ints, o_reals, f"ccd {name}", exact=False)
    ok = close(gpu.per_pair, o_pp)
    assert ok.all(), f"ccd {name}: per-pair scale/min_time differ at {np.argwhere(~ok)[:5]}"


def test_ccd_long_interval(ctx):
    """A 4x longer interval (max_time = 4 dt) reaches more leaf pairs and toc values."""
    s = dense_scene()
    rec = s.records
    load(ctx, s, rec)
    mt = np.full(s.pairs.shape[0], 4 * scenes.DT)
    gpu = d2.compare_trees_full_continuous(ctx, s.pairs, mt)
    o_off, o_ints, o_reals, o_pp = oracle_scene(s, rec).ccd(s.pairs, mt)
    assert_same(gpu, o_off, o_ints, o_reals, "ccd dense 4dt", exact=False)
    assert close(gpu.per_pair, o_pp).all()
    assert (o_ints[:, 3] == 3).any() and (o_ints[:, 3] == 1).any()  # edge and vertex contacts present


def test_project_future(ctx):
    s = dense_scene()
    load(ctx, s, s.records)
    steps = np.linspace(1e-3, 0.05, s.n_bodies)
    ctx.project_future(steps)
    got = ctx.download_bodies()
    want = with_future(s.records, steps)
    # sin/cos of the device and of glibc may differ in the last bit
    assert np.allclose(got[:, 14:28], want[:, 14:28], rtol=0, atol=1e-14)


def test_cluster_steps(ctx):
    s = dense_scene()
    rec = s.records
    load(ctx, s, rec)
    pairs = s.pairs
    # two clusters: connected components would be the caller's job; any split works
    cl = (pairs[:, 0] % 2).astype(np.int32)
    mt = np.minimum(rec[pairs[:, 0], 39], rec[pairs[:, 1], 39])
    ctx.upload_pairs(pairs, mt)
    ctx.ccd_run()
    got = ctx.cluster_steps(cl, 2, scenes.DT)
    osc = oracle_scene(s, rec)
    want = [osc.cluster_step(pairs[cl == k], np.float32(scenes.DT)) for k in range(2)]
    assert np.allclose(got, want, rtol=1e-12, atol=0), (got, want)


def test_reference_bucket_fixture(ctx):
    """comparison_dist.h bucket test (test_comparison.cpp:90-128) through the GPU."""
    from test_oracle_kat import PLANE_F, PLANE_V, PYRAMID_F, PYRAMID_V, body_record
    rng = d2.pcg_seed(42, 54)
    trees = [d2.tree_build(PYRAMID_V, PYRAMID_F, rng), d2.tree_build(PLANE_V, PLANE_F, rng)]
    bodies = np.stack([body_record(fut_trans=(0, 0, 1.0), static=True), body_record(fut_trans=(0, 0, 1.105))])
    ctx.upload_trees(trees)
    ctx.upload_bodies([0, 1], bodies)
    assert d2.compare_trees_full(ctx, [[0, 1]], filter=False).n == 8
    assert d2.compare_trees_full(ctx, [[0, 1]], filter=True).n == 1
    # test_continuous_comparison.cpp:222-263 : toc 0.9
    bodies = np.stack([body_record(trans=(0, 0, 1.0), static=True), body_record(trans=(0, 0, 2.0), vel=(0, 0, -1.0))])
    ctx.upload_bodies([0, 1], bodies)
    r = d2.compare_trees_full_continuous(ctx, [[0, 1]], [1.0])
    tocs = r.reals[:, 10]
    assert r.n > 0 and np.isclose(tocs.min(), 0.9, rtol=1.2e-5)


def test_empty_pair_list(ctx):
    s = c1_scene()
    load(ctx, s, s.records)
    r = d2.compare_trees_full(ctx, np.zeros((0, 2), np.int32))
    assert r.n == 0 and r.offsets.tolist() == [0]
    r = d2.compare_trees_full_continuous(ctx, np.zeros((0, 2), np.int32), np.zeros(0))
    assert r.n == 0


def test_bad_arguments_raise(ctx):
    s = c1_scene()
    load(ctx, s, s.records)
    with pytest.raises(d2.D2Error):
        ctx.upload_pairs([[0, s.n_bodies]])  # body index out of range
    with pytest.raises(d2.D2Error):
        ctx.upload_pairs([[0, 1]], [-1.0])  # negative interval


def test_capacity_error_is_reported():
    s = dense_scene.__wrapped__(seed=8)
    with d2.Context(0) as c:
        c.set_capacity(1024)
        load(c, s, with_future(s.records, 4 * scenes.DT))
        with pytest.raises(d2.CapacityError):
            d2.compare_trees_full(c, s.pairs, filter=False)


@pytest.mark.parametrize("graph", [False, True])
def test_capacities_grow_on_demand(ctx, graph):
    """From tiny capacities with growth on, narrow and CCD runs repeat with
    larger lists and return exactly the default-capacity results."""
    s = dense_scene()
    rec = with_future(s.records, scenes.DT)
    load(ctx, s, rec)
    mt = np.full(s.pairs.shape[0], 4 * scenes.DT)
    ref_n = d2.compare_trees_full(ctx, s.pairs, filter=True)
    ref_c = d2.compare_trees_full_continuous(ctx, s.pairs, mt)
    with d2.Context(0) as c:
        c.set_capacity(1024)
        c.set_auto_grow(True)
        c.set_graph(graph)
        load(c, s, rec)
        for _ in range(2):
            got = d2.compare_trees_full(c, s.pairs, filter=True)
            assert np.array_equal(got.ints, ref_n.ints) and np.array_equal(got.reals, ref_n.reals)
            got = d2.compare_trees_full_continuous(c, s.pairs, mt)
            assert np.array_equal(got.ints, ref_c.ints) and np.array_equal(got.reals, ref_c.reals)
            assert np.array_equal(got.per_pair, ref_c.per_pair)


def test_empty_run_after_contacts_reports_zero(ctx):
    """Regression: an empty pair list right after a run with contacts must not
    reuse the previous run's contact count (it once overflowed host buffers)."""
    s = dense_scene()
    load(ctx, s, s.records)
    mt = np.full(s.pairs.shape[0], scenes.DT)
    assert ctx.upload_pairs(s.pairs, mt) is None and ctx.ccd_run() > 0
    for run in (lambda: ctx.ccd_run(), lambda: ctx.narrow_run(True), lambda: ctx.narrow_run(False)):
        ctx.upload_pairs(np.zeros((0, 2), np.int32), np.zeros(0))
        assert run() == 0
        r = ctx.result()
        assert r.n == 0 and r.ints.shape == (0, 8) and r.offsets.tolist() == [0]


@pytest.fixture(scope="module")
def ctx32():
    c = d2.Context(0, d2.D2_FP32)
    yield c
    c.close()


@pytest.mark.parametrize("name,make", SCENES)
@pytest.mark.parametrize("filt", [False, True])
def test_narrow_parity_fp32_cull(ctx32, name, make, filt):
    """FP32 mode: conservative FP32 culling, FP64 for every kept test -- the
    contacts must be the oracle's exactly (same bar as FP64)."""
    s = make()
    print(f"[narrow fp32] scene={name} pairs={s.pairs.shape[0]} filter={filt}", flush=True)
    rec = with_future(s.records, scenes.DT)
    load(ctx32, s, rec)
    ctx32.reset_stats()
    ctx32.set_profiling(True)
    gpu = d2.compare_trees_full(ctx32, s.pairs, filter=filt)
    ctx32.set_profiling(False)
    o_off, o_ints, o_reals = oracle_scene(s, rec).narrow(s.pairs, filter=filt)
    assert_same(gpu, o_off, o_ints, o_reals, f"narrow fp32 {name} filter={filt}")
    st = ctx32.kernel_stats()
    tested = st["tri_dist"][2]
    # node pairs seen by the mask kernel x up to 64 triangle pairs: the FP64
    # work after the cull must be a small part of the full cross product
    assert tested < 0.5 * 64 * st["tri_enum"][2], (tested, st["tri_enum"])


@pytest.mark.parametrize("name,make", SCENES)
def test_ccd_parity_fp32_culls(ctx32, name, make):
    """FP32 culls of the CCD leaf sub-tests keep the FP64 result exactly and
    drop most sub-tests (a 4-step interval reaches many leaf pairs)."""
    s = make()
    rec = s.records
    load(ctx32, s, rec)
    mt = np.full(s.pairs.shape[0], 4 * scenes.DT)
    ctx32.reset_stats()
    ctx32.set_profiling(True)
    gpu = d2.compare_trees_full_continuous(ctx32, s.pairs, mt)
    ctx32.set_profiling(False)
    o_off, o_ints, o_reals, o_pp = oracle_scene(s, rec).ccd(s.pairs, mt)
    assert_same(gpu, o_off, o_ints, o_reals, f"ccd fp32 {name}", exact=False)
    assert close(gpu.per_pair, o_pp).all()
    st = ctx32.kernel_stats()
    leaf_pairs = int(sum(lv[:, 1].sum() for lv in [ctx32.run_info()[1][1]]))
    kept = st["leaf_point"][2] + st["leaf_edge"][2]
    print(f"[ccd32] {name}: leaf node pairs {leaf_pairs}, sub-tests kept {kept:.0f}", flush=True)
    if name == "dense":
        full = 15 * 64 * leaf_pairs  # at most 6nm + 9nm sub-tests per leaf pair (n, m <= 8)
        assert kept < 0.5 * full
