"""GPU tests of the learned lambda warm start (no reference counterpart):
training data generated by the solver itself, training with the library's
own BF16 MFMA kernels, and what the trained network gains over the zero
start at the same iteration count."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _cloth(res=32):
    from xpbd_amd import Solver
    from xpbd_amd.scenes import cloth_sheet

    return cloth_sheet(Solver, res=res, solve_mode="jacobi", device=DEV)[0]


@pytest.fixture(scope="module")
def trained():
    from xpbd_amd import warmstart as ws

    log = []
    theta = ws.fit(_cloth, substeps=120, settle_frames=2, steps=800, log=log)
    return theta, log[0]


def test_training_lowers_the_loss_and_data_stay_finite(trained):
    """Data generation (plain, then on-policy) and training on the GPU: the
    final loss is finite and the on-policy round's states stayed finite."""
    import torch

    from xpbd_amd import warmstart as ws

    theta, log = trained
    print("fit log", log)
    assert log["round2_data_finite"]
    assert bool(torch.isfinite(theta).all())
    assert all(np.isfinite(v) for _, v in log["round1_loss"] + log["round2_loss"])
    # the trained network fits fresh plain-trajectory data better than the
    # initial parameters do (targets are logits of the first-step multiple)
    F, T = ws.collect(_cloth(), 30, every=3)
    y0 = ws.predict(torch.from_numpy(ws.init_params(0)).to(F.device), F)
    y1 = ws.predict(theta, F)
    e0, e1 = float(((y0 - T) ** 2).mean()), float(((y1 - T) ** 2).mean())
    print(f"held-out MSE: initial {e0:.3f}, trained {e1:.3f}")
    assert e1 < 0.9 * e0


def test_learned_start_beats_zero_start_at_equal_iterations(trained):
    """On a fresh run of the same scene, each start on its own trajectory:
    RMS residual after k iterations from the learned start against k (and
    2k) iterations from zero."""
    from xpbd_amd import warmstart as ws

    theta, _ = trained
    s = _cloth()
    for _ in range(20):
        s.step()
    res, fallback = ws.Evaluator(s, theta).run(substeps=60, plain_iters=(1, 2, 4),
                                               learned_iters=(1, 2))
    print("rms residual", res, "fallbacks", fallback)
    assert res["learned_1"] < res["plain_1"]
    assert res["learned_2"] < res["plain_2"]
    # reaches the zero start's residual with half the iterations
    assert res["learned_2"] <= res["plain_4"]


def test_frame_plan_with_warm_start_and_guard(trained):
    """The warm start inside the captured frame plan runs and stays finite;
    the guard's first check records both residuals and picks the better."""
    from xpbd_amd import warmstart as ws

    theta, _ = trained
    s = _cloth()
    g = ws.Guard(s, theta, check_every=5)
    g.step(12)
    assert len(g.history) == 3
    for _, rp, rl, on in g.history:
        assert on == (rl <= rp)
    assert np.all(np.isfinite(s.positions()))
    s.set_warm_start(theta)
    s.step(30)
    assert np.all(np.isfinite(s.positions()))
    s.set_warm_start(None)
    s.step(2)
    assert np.all(np.isfinite(s.positions()))


def test_ws_kernels_match_numpy_oracle():
    """xpbd_ws_features / xpbd_ws_apply / xpbd_ws_residual against the
    float64 NumPy restatement (oracle/ws_ref.py) on a cloth with random
    velocities, part of it pushed below the floor, and random network
    outputs."""
    import torch

    from oracle import ws_ref
    from xpbd_amd import _native as N
    from xpbd_amd import warmstart as ws

    s = _cloth(16)
    ws._prepare(s)
    rng = np.random.default_rng(3)
    X = s.X.cpu().numpy().copy()
    X[:, 1] -= 0.5 + rng.normal(0, 2e-3, len(X)).astype(np.float32)   # about half below y = 0
    X[:, :3] += rng.normal(0, 1e-3, (len(X), 3)).astype(np.float32)
    V = np.zeros_like(X)
    V[:, :3] = rng.normal(0, 0.3, (len(X), 3)).astype(np.float32)
    s.X.copy_(torch.from_numpy(X))
    s.V.copy_(torch.from_numpy(V))
    m = s.num_constraints
    h = float(np.float32(s.dt / s.substeps))
    lib = N.lib()
    F = torch.empty((m, 12), device=DEV)
    N.check(lib.xpbd_ws_features(None, N.ptr(s.X), N.ptr(s.V), N.ptr(s.R), N.ptr(s.ANC), None, h,
                                 N.ptr(F), m))
    rows = s.R[:m].cpu().numpy()
    anc = s.ANC[:m].cpu().numpy()
    Fr = ws_ref.features(X, V, rows, anc, h)
    assert np.allclose(F.cpu().numpy(), Fr, rtol=1e-5, atol=2e-5)
    Y = torch.from_numpy(rng.normal(0, 2, m).astype(np.float32)).to(DEV)
    LAM = torch.zeros(m, device=DEV)
    ACC = torch.zeros_like(s.X)
    N.check(lib.xpbd_ws_apply(None, N.ptr(s.X), N.ptr(s.V), N.ptr(s.R), N.ptr(s.ANC), N.ptr(Y), h,
                              N.ptr(LAM), N.ptr(ACC), m))
    lam_r, acc_r = ws_ref.apply(X, V, rows, anc, Y.cpu().numpy(), h)
    lam_g = LAM.cpu().numpy()
    assert (lam_g != 0).sum() > m // 4
    assert np.allclose(lam_g, lam_r, rtol=1e-4, atol=1e-10)
    acc_g = ACC.cpu().numpy()
    assert np.array_equal(acc_g[:, 3], acc_r[:, 3])
    # float atomic sums of terms of both signs: compare against the scale of
    # the largest sum
    err = np.abs(acc_g[:, :3] - acc_r[:, :3]).max()
    print(f"offset sums: max |gpu - oracle| {err:.3e}, max |oracle| {np.abs(acc_r[:, :3]).max():.3e}")
    assert err <= 1e-5 * np.abs(acc_r[:, :3]).max()
    out = torch.zeros(1, dtype=torch.float64, device=DEV)
    N.check(lib.xpbd_ws_residual(None, N.ptr(s.X), N.ptr(s.R), N.ptr(s.ANC), N.ptr(LAM), h,
                                 N.ptr(out), m))
    r_ref = ws_ref.residual_sum(X, rows, anc, lam_g, h)
    assert np.isclose(float(out.item()), r_ref, rtol=1e-4)
