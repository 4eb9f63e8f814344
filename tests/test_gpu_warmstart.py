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
    # initial parameters do
    F, T = ws.collect(_cloth(), 30, every=3)
    y0 = ws.predict(torch.from_numpy(ws.init_params(0)).to(F.device), F)
    y1 = ws.predict(theta, F)
    assert float(((y1 - T) ** 2).mean()) < 0.5 * float(((y0 - T) ** 2).mean())


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
