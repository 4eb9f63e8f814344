"""The warm-start MLP: the NumPy oracle (oracle/mlp_ref.py) pinned by
central finite differences and the Adam closed form on CPU, and the HIP
MFMA forward / gradient / Adam kernels against the oracle on the GPU."""
import numpy as np
import pytest


def test_param_count_and_layout():
    from oracle import mlp_ref as M

    assert M.N_PARAMS == 5057
    from xpbd_amd import _native as N

    assert N.lib().xpbd_mlp_param_count() == 5057


def test_oracle_gradient_matches_finite_differences():
    from oracle import mlp_ref as M

    rng = np.random.default_rng(0)
    theta = M.init_params(1).astype(np.float64)
    X = rng.normal(size=(20, 12))
    t = rng.normal(size=20)
    _, g = M.loss_and_grad(theta, X, t, bf16_emulate=False)
    idx = rng.choice(M.N_PARAMS, 60, replace=False)
    idx = np.concatenate([idx, [M.N_PARAMS - 1, 0, 768, 768 + 64 + 5]])
    for i in idx:
        e = np.zeros_like(theta)
        e[i] = 1e-6
        lp, _ = M.loss_and_grad(theta + e, X, t, bf16_emulate=False)
        lm, _ = M.loss_and_grad(theta - e, X, t, bf16_emulate=False)
        fd = (lp - lm) / 2e-6
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_oracle_gelu_and_bf16():
    from oracle import mlp_ref as M

    z = np.array([-3.0, -1.0, 0.0, 0.5, 2.0])
    # GELU(0) = 0, GELU(x) - GELU(-x) = x, GELU'(0) = 1/2
    assert M.gelu(np.array([0.0]))[0] == 0.0
    assert np.allclose(M.gelu(z) - M.gelu(-z), z)
    assert M.gelu_grad(np.array([0.0]))[0] == pytest.approx(0.5)
    # the tanh form stays within 1e-3 of the erf definition (its published
    # worst case is ~5e-4 near |z| = 2); the derivative by central differences
    g = np.linspace(-6.0, 6.0, 2401)
    assert np.abs(M.gelu(g) - M.gelu_exact(g)).max() < 1e-3
    fd = (M.gelu(g + 1e-6) - M.gelu(g - 1e-6)) / 2e-6
    assert np.allclose(M.gelu_grad(g), fd, atol=1e-7)
    # bf16 keeps 8 significant bits, ties to even
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 3.14159265], np.float32)
    assert M.bf16(x).tolist() == [1.0, 1.0, 1.0 + 2 ** -6, 3.140625]


def test_oracle_adam_first_step_is_signed_lr():
    """With zero moments, the first Adam step moves each parameter by
    lr * g / (|g| + eps) (bias corrections cancel the (1-beta) factors)."""
    from oracle import mlp_ref as M

    th = np.zeros(5, np.float32)
    m = np.zeros(5, np.float32)
    v = np.zeros(5, np.float32)
    g = np.array([1.0, -2.0, 0.5, 0.0, 3.0], np.float32)
    M.adam(th, m, v, g, 1, lr=0.01)
    assert np.allclose(th, -0.01 * np.sign(g), atol=1e-6)


@pytest.mark.gpu
def test_mlp_forward_grad_adam_match_oracle():
    import torch

    from oracle import mlp_ref as M
    from xpbd_amd import _native as N

    lib = N.lib()
    rng = np.random.default_rng(0)
    theta = M.init_params(3)
    m = 1000
    X = rng.normal(size=(m, 12)).astype(np.float32)
    t = rng.normal(size=m).astype(np.float32)
    dev = "cuda:0"
    th = torch.from_numpy(theta).to(dev)
    F = torch.from_numpy(X).to(dev)
    T = torch.from_numpy(t).to(dev)
    Y = torch.zeros(m, device=dev)
    N.check(lib.xpbd_mlp_forward(N.current_stream(), N.ptr(th), N.ptr(F), N.ptr(Y), m))
    y_ref, _ = M.forward(theta, X, bf16_emulate=True)
    y = Y.cpu().numpy()
    assert np.abs(y - y_ref).max() <= 1e-3 * max(1.0, np.abs(y_ref).max())
    nb = 8
    ws = torch.empty(int(lib.xpbd_mlp_train_workspace_bytes(nb)), dtype=torch.uint8, device=dev)
    G = torch.zeros(5057, device=dev)
    L = torch.zeros(1, device=dev)
    N.check(lib.xpbd_mlp_grad(N.current_stream(), N.ptr(th), N.ptr(F), N.ptr(T), m, nb,
                              N.ptr(ws), N.ptr(G), N.ptr(L)))
    loss_ref, g_ref = M.loss_and_grad(theta, X, t, bf16_emulate=True)
    g = G.cpu().numpy()
    assert float(L.item()) == pytest.approx(loss_ref, rel=1e-3)
    p, q = M.unpack(g), M.unpack(g_ref)
    for name in ("W1", "b1", "W2", "b2", "W3", "b3"):
        err = np.abs(p[name] - q[name]).max() / max(np.abs(q[name]).max(), 1e-12)
        assert err < 2e-2, (name, err)   # bf16 rounding of the backward operands
    mom = torch.zeros(5057, device=dev)
    vel = torch.zeros(5057, device=dev)
    N.check(lib.xpbd_mlp_adam(N.current_stream(), N.ptr(th), N.ptr(mom), N.ptr(vel), N.ptr(G),
                              1e-3, 0.9, 0.999, 1e-8, 1))
    th_ref = theta.copy()
    M.adam(th_ref, np.zeros(5057, np.float32), np.zeros(5057, np.float32), g, 1, lr=1e-3)
    assert np.abs(th.cpu().numpy() - th_ref).max() < 1e-6
