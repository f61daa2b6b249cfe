"""GPU: the learned pair gate.

- The training kernels' loss and gradient equal the CPU oracle
  (oracle/mlp_oracle.py, itself pinned by finite differences).
- The MFMA inference kernel equals the oracle with bf16 activations.
- Device training lowers the loss on labels from plain CCD runs.
- With the gate on, every run returns exactly the plain path's result; a
  network that proposes every pair leaves the certificate alone to decide,
  so its soundness is tested on every pair of the scenes.
- The trained gate certifies pairs and removes their traversal work.
"""
import numpy as np
import pytest

import delta2_amd as d2
import mlp_oracle as mo
from delta2_amd import learned, scenes
from gpu_cases import c1_scene, dense_scene, hull_scene, sparse_scene, with_future

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = d2.Context(0, d2.D2_FP64)
    yield c
    c.close()


def propose_all():
    """Zero weights, b3 = (1, 2): every pair is proposed as contact-free."""
    p = np.zeros(mo.N_PARAMS, dtype=np.float32)
    p[-2:] = (1.0, 2.0)
    return p


def test_training_gradient_matches_oracle(ctx):
    rng = np.random.default_rng(3)
    X = rng.normal(0, 1, (700, 24)).astype(np.float32)
    T = rng.normal(0, 0.5, (700, 2)).astype(np.float32)
    parts = mo.unpack(mo.init_params(5))
    parts["b1"] = parts["b1"] + 0.05
    params = mo.pack(parts).astype(np.float32)
    ctx.mlp_set_params(params)
    loss, g = ctx.mlp_loss_grad(X, T, lw=(1.0, 0.5), delta=0.3)
    ol, og = mo.loss_grad(params, X, T, lw=(1.0, 0.5), delta=0.3)
    print(f"[mlp] loss gpu {loss:.6f} oracle {ol:.6f}, grad max abs diff {np.abs(g - og).max():.3e}", flush=True)
    assert abs(loss - ol) <= 1e-5 * max(1.0, abs(ol))
    assert np.allclose(g, og, rtol=1e-3, atol=1e-6)
    assert np.array_equal(ctx.mlp_get_params(), params)


def test_mfma_inference_matches_oracle(ctx):
    rng = np.random.default_rng(4)
    X = rng.normal(0, 1, (1000 + 7, 24)).astype(np.float32)  # a ragged last tile
    params = mo.init_params(6)
    params[-2:] = (0.3, -0.2)
    ctx.mlp_set_params(params)
    Y = ctx.mlp_infer(X)
    Yo = mo.forward(params, X, bf16_act=True)
    Yf = mo.forward(params, X, bf16_act=False)
    err = np.abs(Y - Yo)
    print(f"[mlp] infer max err {err.max():.3e} median {np.median(err):.3e}, vs fp32 path {np.abs(Y - Yf).max():.3e}",
          flush=True)
    # FP32 sums in another order may round a hidden activation to the next bf16 value
    assert err.max() <= 2e-2 * max(1.0, np.abs(Yo).max())
    assert np.median(err) <= 1e-3
    assert np.abs(Y - Yf).max() <= 0.1 * max(1.0, np.abs(Yf).max())


def _dataset(ctx):
    Xs, Ts = [], []
    for make in (c1_scene, dense_scene, hull_scene, sparse_scene):
        s = make()
        ctx.upload_trees(s.trees)
        ctx.upload_bodies(s.tree_of_body, s.records)
        mt = np.minimum(s.records[s.pairs[:, 0], 39], s.records[s.pairs[:, 1], 39])
        ctx.upload_pairs(s.pairs, mt)
        ctx.ccd_run()
        Ts.append(learned.labels(ctx.result(with_pairs=True), s.pairs.shape[0]))
        Xs.append(ctx.pair_features())
    return np.concatenate(Xs), np.concatenate(Ts)


def test_device_training_lowers_loss(ctx):
    X, T = _dataset(ctx)
    assert np.isfinite(X).all() and X.shape[1] == 24
    params, losses = learned.train(ctx, X, T, epochs=30, batch=256, lr=3e-3)
    print(f"[mlp] {X.shape[0]} pairs, loss {losses[0]:.4f} -> {losses[-1]:.4f}", flush=True)
    assert losses[-1] < 0.5 * losses[0]
    loss, _ = ctx.mlp_loss_grad(X, T, learned.LW, learned.DELTA)
    assert abs(loss - losses[-1]) < 0.5 * losses[0]


def _runs(c, s, rec, learned_on, params=None):
    """Narrow, CCD and step results of one context on one scene."""
    if params is not None:
        c.mlp_set_params(params)
    c.set_learned(learned_on)
    c.upload_trees(s.trees)
    c.upload_bodies(s.tree_of_body, rec)
    pairs = s.pairs
    mt = np.minimum(rec[pairs[:, 0], 39], rec[pairs[:, 1], 39])
    c.upload_pairs(pairs, mt)
    c.narrow_run(True)
    out = {"narrow": c.result()}
    c.ccd_run()
    out["ccd"] = c.result(with_pairs=True)
    out["ccd_gate"] = c.gate_counts()["ccd"]
    cl, k = scenes.clusters(pairs, s.n_bodies, rec)
    c.upload_pairs(pairs)
    c.upload_clusters(cl, k)
    _, steps = c.step_run(scenes.DT, True, want_steps=True)
    out["step"] = c.result(with_pairs=True)
    out["steps"] = steps
    out["step_gate"] = c.gate_counts()
    g, lv = c.run_info()
    out["step_node_pairs"] = int(lv[0, :, 0].sum() + lv[1, :, 0].sum() + lv[1, :, 1].sum())
    return out


def _same(a, b, label):
    for k in ("narrow", "ccd", "step"):
        x, y = a[k], b[k]
        assert np.array_equal(x.offsets, y.offsets), f"{label} {k}: offsets"
        assert np.array_equal(x.ints, y.ints), f"{label} {k}: index tuples"
        assert np.array_equal(x.reals, y.reals), f"{label} {k}: reals"
        if x.per_pair is not None:
            assert np.array_equal(x.per_pair, y.per_pair), f"{label} {k}: per-pair values"
    assert np.array_equal(a["steps"], b["steps"]), f"{label}: cluster steps"


@pytest.mark.parametrize("precision", [d2.D2_FP64, d2.D2_FP32])
@pytest.mark.parametrize("name,make", [("sparse", sparse_scene), ("dense", dense_scene), ("hulls", hull_scene)])
def test_certificate_alone_keeps_results(precision, name, make):
    s = make()
    rec = with_future(s.records, scenes.DT)
    print(f"[gate] scene={name} precision={precision} pairs={s.pairs.shape[0]}", flush=True)
    with d2.Context(0, precision) as plain, d2.Context(0, precision) as gated:
        a = _runs(plain, s, rec, False)
        b = _runs(gated, s, rec, True, params=propose_all())
        _same(a, b, f"{name}/{precision}")
        print(f"[gate] proposed/certified ccd {b['ccd_gate']} step {b['step_gate']}", flush=True)
        assert b["ccd_gate"][0] == s.pairs.shape[0]  # every pair proposed
        if name == "sparse":
            assert b["step_gate"]["ccd"][1] > 0 and b["step_gate"]["discrete"][1] > 0
            assert b["step_node_pairs"] < a["step_node_pairs"]
        # a certified pair has no contact in the plain result
        cert_free = np.diff(a["ccd"].offsets) == 0
        assert b["ccd_gate"][1] <= int(cert_free.sum())


def test_trained_gate_removes_work_and_keeps_results(ctx):
    X, T = _dataset(ctx)
    params, _ = learned.train(ctx, X, T, epochs=30, batch=256, lr=3e-3)
    s = sparse_scene.__wrapped__(seed=1303, n=256)  # held out
    rec = with_future(s.records, scenes.DT)
    with d2.Context(0, d2.D2_FP32) as plain, d2.Context(0, d2.D2_FP32) as gated:
        a = _runs(plain, s, rec, False)
        b = _runs(gated, s, rec, True, params=params)
        _same(a, b, "held-out")
        stats = learned.gate_stats(gated.mlp_infer(gated.pair_features()), learned.labels(a["ccd"], s.pairs.shape[0]))
        print(f"[gate] trained: {b['step_gate']}, node pairs {a['step_node_pairs']} -> {b['step_node_pairs']}, "
              f"{stats}", flush=True)
        assert b["step_gate"]["ccd"][1] > 0
        assert b["step_node_pairs"] < a["step_node_pairs"]
