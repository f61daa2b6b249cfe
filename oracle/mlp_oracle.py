"""CPU oracle of the learned pair gate's network (test infrastructure only).

The 24 -> 64 -> 64 -> 2 MLP of the library (ReLU, bf16 weights, FP32 master
weights), restated in numpy float64 for checking the HIP kernels:
  * forward(params, X, bf16_act): the network on bf16-rounded weights; with
    bf16_act the inputs and hidden activations are rounded to bf16 as the
    MFMA inference kernel does;
  * loss_grad(params, X, T, lw, delta): the training loss (mean over samples of
    sum_o lw_o * huber_delta(y_o - t_o)) and its gradient with respect to the
    (bf16-rounded) weights, which the training kernels apply to the FP32
    master weights.
The gradient is pinned by central finite differences (tests/test_mlp_oracle.py).
The network has no counterpart in the reference; its layout is the library's:
W1[24][64] b1[64] W2[64][64] b2[64] W3[64][2] b3[2], row = input.
"""
import numpy as np

N_IN, N_H, N_OUT = 24, 64, 2
SIZES = [("W1", (N_IN, N_H)), ("b1", (N_H,)), ("W2", (N_H, N_H)), ("b2", (N_H,)), ("W3", (N_H, N_OUT)),
         ("b3", (N_OUT,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SIZES)


def bf16(x):
    """Round float32 values to bf16 (nearest, ties to even), returned as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return (u & 0xFFFFFFFF).astype(np.uint32).view(np.float32)


def unpack(params):
    p = np.asarray(params, dtype=np.float64).ravel()
    assert p.size == N_PARAMS
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = p[o:o + n].reshape(shape)
        o += n
    return out


def pack(parts):
    return np.concatenate([np.asarray(parts[name], dtype=np.float64).ravel() for name, _ in SIZES])


def init_params(seed=0):
    """He-normal weights, zero biases."""
    rng = np.random.default_rng(seed)
    parts = {"W1": rng.normal(0, np.sqrt(2.0 / N_IN), (N_IN, N_H)), "b1": np.zeros(N_H),
             "W2": rng.normal(0, np.sqrt(2.0 / N_H), (N_H, N_H)), "b2": np.zeros(N_H),
             "W3": rng.normal(0, np.sqrt(1.0 / N_H), (N_H, N_OUT)), "b3": np.zeros(N_OUT)}
    return pack(parts).astype(np.float32)


def effective(params):
    """The weights the kernels compute with: bf16-rounded, as float64."""
    return unpack(bf16(np.asarray(params, dtype=np.float32)).astype(np.float64))


def forward(params, X, bf16_act=False, w=None):
    w = effective(params) if w is None else w
    X = np.asarray(X, dtype=np.float64)
    if bf16_act:
        X = bf16(X.astype(np.float32)).astype(np.float64)
    h1 = np.maximum(X @ w["W1"] + w["b1"], 0.0)
    if bf16_act:
        h1 = bf16(h1.astype(np.float32)).astype(np.float64)
    h2 = np.maximum(h1 @ w["W2"] + w["b2"], 0.0)
    if bf16_act:
        h2 = bf16(h2.astype(np.float32)).astype(np.float64)
    return h2 @ w["W3"] + w["b3"]


def loss_w(w, X, T, lw=(1.0, 1.0), delta=0.1):
    """Loss as a function of the effective weights (used for finite differences)."""
    y = forward(None, X, w=w)
    r = y - np.asarray(T, dtype=np.float64)
    a = np.abs(r)
    h = np.where(a <= delta, 0.5 * r * r, delta * (a - 0.5 * delta))
    return float(np.mean(h @ np.asarray(lw, dtype=np.float64)))


def loss_grad(params, X, T, lw=(1.0, 1.0), delta=0.1, w=None):
    w = effective(params) if w is None else w
    X = np.asarray(X, dtype=np.float64)
    T = np.asarray(T, dtype=np.float64)
    n = X.shape[0]
    z1 = X @ w["W1"] + w["b1"]
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ w["W2"] + w["b2"]
    h2 = np.maximum(z2, 0.0)
    y = h2 @ w["W3"] + w["b3"]
    r = y - T
    a = np.abs(r)
    lwv = np.asarray(lw, dtype=np.float64)
    loss = float(np.mean(np.where(a <= delta, 0.5 * r * r, delta * (a - 0.5 * delta)) @ lwv))
    d3 = lwv * np.clip(r, -delta, delta) / n
    g = {"W3": h2.T @ d3, "b3": d3.sum(0)}
    d2 = (d3 @ w["W3"].T) * (z2 > 0)
    g["W2"] = h1.T @ d2
    g["b2"] = d2.sum(0)
    d1 = (d2 @ w["W2"].T) * (z1 > 0)
    g["W1"] = X.T @ d1
    g["b1"] = d1.sum(0)
    return loss, pack(g)
