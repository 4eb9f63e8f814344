"""ORACLE — test infrastructure only.  NumPy reference of the warm-start MLP
(12 -> 64 -> 64 -> 1, GELU in its tanh form) used to check the HIP training and
inference kernels.  No reference counterpart: the learned warm start is new.

Mixed precision is emulated at the points where the device kernels round:
the inputs and first-layer weights, the first hidden activations and the
second-layer weights are rounded to bfloat16 (round to nearest even) before
the two 64-wide products, which accumulate in float32; the output layer runs
in float32.  ``forward``/``backward`` with ``bf16=False`` are the plain
float64 network, pinned against central finite differences in
tests/test_mlp.py.
"""
from __future__ import annotations

import math

import numpy as np

N_IN, N_H, N_OUT = 12, 64, 1
SHAPES = [("W1", (N_IN, N_H)), ("b1", (N_H,)), ("W2", (N_H, N_H)), ("b2", (N_H,)),
          ("W3", (N_H,)), ("b3", (1,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)   # 5057


def unpack(theta: np.ndarray) -> dict:
    out, o = {}, 0
    for name, shape in SHAPES:
        k = int(np.prod(shape))
        out[name] = theta[o:o + k].reshape(shape)
        o += k
    return out


def bf16(x: np.ndarray) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest, ties to even), returned as float32."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u >> 16) & 1) + 0x7FFF
    y = ((u + r) >> 16 << 16).astype(np.uint32)
    return y.view(np.float32)


_erf = np.vectorize(math.erf, otypes=[np.float64])
C0 = 2.0 * math.sqrt(2.0 / math.pi)
C1 = 0.044715


def gelu_exact(z):
    """0.5 z (1 + erf(z / sqrt 2)), the definition the tanh form approximates."""
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def _sig(z):
    return 1.0 / (1.0 + np.exp(-C0 * z * (1.0 + C1 * z * z)))


def gelu(z):
    """Tanh form: 0.5 z (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3))) = z sigmoid(u)."""
    return z * _sig(z)


def gelu_grad(z):
    s = _sig(z)
    return s + z * s * (1.0 - s) * C0 * (1.0 + 3.0 * C1 * z * z)


def init_params(seed: int = 0) -> np.ndarray:
    """He-style initialisation, float32."""
    rng = np.random.default_rng(seed)
    theta = np.zeros(N_PARAMS, np.float32)
    p = unpack(theta)
    p["W1"][:] = rng.normal(0, math.sqrt(2.0 / N_IN), p["W1"].shape)
    p["W2"][:] = rng.normal(0, math.sqrt(2.0 / N_H), p["W2"].shape)
    p["W3"][:] = rng.normal(0, math.sqrt(1.0 / N_H), p["W3"].shape)
    return theta


def forward(theta, X, bf16_emulate=True):
    """Returns y (N,) and the cache for backward."""
    p = unpack(np.asarray(theta, np.float32 if bf16_emulate else np.float64))
    if bf16_emulate:
        Xq, W1q = bf16(X), bf16(p["W1"])
        Z1 = (Xq.astype(np.float64) @ W1q.astype(np.float64)).astype(np.float32) + p["b1"]
        H1 = gelu(Z1.astype(np.float64)).astype(np.float32)
        H1q, W2q = bf16(H1), bf16(p["W2"])
        Z2 = (H1q.astype(np.float64) @ W2q.astype(np.float64)).astype(np.float32) + p["b2"]
        H2 = gelu(Z2.astype(np.float64)).astype(np.float32)
        y = (H2.astype(np.float64) @ p["W3"].astype(np.float64)).astype(np.float32) + p["b3"][0]
        return y, dict(X=Xq, Z1=Z1, H1=H1q, Z2=Z2, H2=H2, W2=W2q)
    X = np.asarray(X, np.float64)
    Z1 = X @ p["W1"] + p["b1"]
    H1 = gelu(Z1)
    Z2 = H1 @ p["W2"] + p["b2"]
    H2 = gelu(Z2)
    y = H2 @ p["W3"] + p["b3"][0]
    return y, dict(X=X, Z1=Z1, H1=H1, Z2=Z2, H2=H2, W2=p["W2"])


def loss_and_grad(theta, X, t, bf16_emulate=True):
    """Mean squared error and its gradient w.r.t. the flat parameters."""
    theta_f = np.asarray(theta, np.float32 if bf16_emulate else np.float64)
    p = unpack(theta_f)
    y, c = forward(theta_f, X, bf16_emulate)
    n = len(y)
    r = (y.astype(np.float64) - np.asarray(t, np.float64))
    loss = float(np.mean(r * r))
    dy = 2.0 * r / n
    g = np.zeros(N_PARAMS, np.float64)
    gp = unpack(g)
    H2 = c["H2"].astype(np.float64)
    gp["W3"][:] = H2.T @ dy
    gp["b3"][:] = dy.sum()
    dZ2 = np.outer(dy, np.asarray(p["W3"], np.float64)) * gelu_grad(c["Z2"].astype(np.float64))
    gp["W2"][:] = c["H1"].astype(np.float64).T @ dZ2
    gp["b2"][:] = dZ2.sum(axis=0)
    dH1 = dZ2 @ np.asarray(c["W2"], np.float64).T
    dZ1 = dH1 * gelu_grad(c["Z1"].astype(np.float64))
    gp["W1"][:] = c["X"].astype(np.float64).T @ dZ1
    gp["b1"][:] = dZ1.sum(axis=0)
    return loss, g


def adam(theta, m, v, g, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015) in float32, in place; step >= 1."""
    g = np.asarray(g, np.float32)
    m[:] = b1 * m + (1 - b1) * g
    v[:] = b2 * v + (1 - b2) * g * g
    mh = m / np.float32(1 - b1 ** step)
    vh = v / np.float32(1 - b2 ** step)
    theta[:] = theta - np.float32(lr) * mh / (np.sqrt(vh) + np.float32(eps))
