"""CPU: the network oracle (oracle/mlp_oracle.py) pinned by finite differences
and by closed forms, before it is used to check the HIP training kernels."""
import numpy as np

import mlp_oracle as mo


def _data(seed, n=64):
    rng = np.random.default_rng(seed)
    X = rng.normal(0, 1, (n, mo.N_IN))
    T = rng.normal(0, 0.5, (n, mo.N_OUT))
    return X, T


def test_gradient_matches_central_differences():
    X, T = _data(1)
    w = mo.effective(mo.init_params(3))
    w["b1"] = w["b1"] + 0.05  # keep most ReLUs away from their kink
    loss, g = mo.loss_grad(None, X, T, lw=(1.0, 0.5), delta=0.3, w=w)
    flat = mo.pack(w)
    rng = np.random.default_rng(7)
    idx = rng.choice(flat.size, 300, replace=False)
    h = 1e-6
    num = []
    for q in idx:
        wp, wm = flat.copy(), flat.copy()
        wp[q] += h
        wm[q] -= h
        num.append((mo.loss_w(mo.unpack(wp), X, T, (1.0, 0.5), 0.3) - mo.loss_w(mo.unpack(wm), X, T, (1.0, 0.5), 0.3))
                   / (2 * h))
    num = np.array(num)
    assert abs(loss - mo.loss_w(w, X, T, (1.0, 0.5), 0.3)) < 1e-12
    assert np.allclose(g[idx], num, rtol=1e-5, atol=1e-8), np.abs(g[idx] - num).max()


def test_closed_forms():
    # bias-only network: y = b3, gradient of a squared residual is its residual
    p = np.zeros(mo.N_PARAMS, dtype=np.float32)
    parts = mo.unpack(p)
    parts["b3"] = np.array([0.25, -0.5])
    X = np.zeros((4, mo.N_IN))
    T = np.zeros((4, mo.N_OUT))
    loss, g = mo.loss_grad(mo.pack(parts).astype(np.float32), X, T, lw=(1.0, 1.0), delta=10.0)
    assert np.isclose(loss, 0.5 * (0.25 ** 2 + 0.5 ** 2))
    gb3 = mo.unpack(g)["b3"]
    assert np.allclose(gb3, [0.25, -0.5])
    # Huber beyond delta: constant slope delta * sign
    loss, g = mo.loss_grad(mo.pack(parts).astype(np.float32), X, T, lw=(2.0, 1.0), delta=0.1)
    assert np.allclose(mo.unpack(g)["b3"], [2.0 * 0.1, -0.1])


def test_bf16_rounding():
    # bf16 keeps 7 mantissa bits: the spacing at 1.0 is 2^-7
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7 + 2 ** -8, 1.0 + 2 ** -7 + 2 ** -9, 1.0 + 2 ** -8 + 2 ** -10,
                  -3.5, 1e-3], dtype=np.float32)
    r = mo.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0  # half way, ties to the even mantissa
    assert r[2] == np.float32(1.0 + 2 ** -6)  # half way from an odd mantissa: up
    assert r[3] == np.float32(1.0 + 2 ** -7)  # below half: down
    assert r[4] == np.float32(1.0 + 2 ** -7)  # above half: up
    assert r[5] == -3.5
    assert abs(r[6] - 1e-3) <= 2 ** -8 * 1e-3
