"""GPU tests of the learned lambda warm start (no reference counterpart):
the warm-start Jacobi launch against the float64 oracle and against the plain
launch, training data generated by the solver itself, training with the
library's own BF16 MFMA kernels, and what the trained network gains over the
zero start."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _cloth(res=32, **kw):
    from xpbd_amd import Solver
    from xpbd_amd.scenes import cloth_sheet

    return cloth_sheet(Solver, res=res, solve_mode="jacobi", device=DEV, **kw)[0]


def _theta(seed=0, b3=None):
    import torch

    from xpbd_amd import warmstart as ws

    th = ws.init_params(seed)
    th[-1] = 0.0 if b3 is None else b3
    return torch.from_numpy(th).to(DEV)


def _warm(s, src, dst, lam, theta, Y=None, predict=0):
    """Network pass + warm-start launch."""
    import torch

    from xpbd_amd import _native as N

    h = float(np.float32(s.dt / s.substeps))
    trs, hs, hr = (N.ptr(t) for t in s._tile_dev)
    if Y is None:
        Y = torch.empty(s.num_constraints, dtype=torch.bfloat16, device=DEV)
    img = torch.empty(int(N.lib().xpbd_mlp_image_bytes()), dtype=torch.uint8, device=DEV)
    N.check(N.lib().xpbd_mlp_pack(None, N.ptr(theta), N.ptr(img)))
    N.check(N.lib().xpbd_ws_net(None, N.ptr(src), N.ptr(s.XP), N.ptr(s.V), N.ptr(s.R),
                                N.ptr(s.ANC), N.ptr(img), N.ptr(Y), h, *s.gravity, None,
                                s.num_constraints, predict, None, 0))
    N.check(N.lib().xpbd_jacobi_tile_warm(
        None, N.ptr(src), N.ptr(dst), N.ptr(s.XP), N.ptr(s.V), N.ptr(s.R), N.ptr(s.ANC),
        N.ptr(lam), trs, hs, hr, N.ptr(Y), h, *s.gravity, None, s._tile, s.num_bodies,
        s.jacobi_relax, predict, None, 0))


def _perturbed_state(s, seed, below_floor=False):
    """Positions jittered, x_prev = x - v h with random velocities."""
    import torch

    rng = np.random.default_rng(seed)
    h = np.float32(s.dt / s.substeps)
    X = s.X.cpu().numpy().copy()
    X[:, :3] += rng.normal(0, 2e-3, (len(X), 3)).astype(np.float32)
    if below_floor:
        X[:, 1] -= 0.5 + rng.normal(0, 2e-3, len(X)).astype(np.float32)
    V = rng.normal(0, 0.3, (len(X), 3)).astype(np.float32)
    XP = X.copy()
    XP[:, :3] = X[:, :3] - V * h
    XP[:, 3] = 0.0
    s.X.copy_(torch.from_numpy(X))
    s.XP.copy_(torch.from_numpy(XP))
    s.V[:, :3] = torch.from_numpy(V).to(DEV)
    return X, XP


def test_warm_launch_matches_oracle():
    """Network pass + warm-start launch on a jittered 24^2 cloth, half of it
    below the floor, with random velocities and a random network: (1) every
    row's output equals xpbd_ws_features -> xpbd_mlp_forward bit for bit;
    (2) those features match the float64 oracle; (3) lambda0 and the moved
    positions match oracle/ws_ref.warm_launch given the same outputs."""
    import torch

    from oracle import ws_ref
    from xpbd_amd import _native as N
    from xpbd_amd import warmstart as ws

    s = _cloth(24)
    ws._prepare(s)
    X, XP = _perturbed_state(s, 5, below_floor=True)
    m = s.num_constraints
    h = float(np.float32(s.dt / s.substeps))
    theta = _theta(7, b3=-1.0)
    Xout = torch.empty_like(s.X)
    lam = torch.zeros(m, device=DEV)
    Y = torch.zeros(m, dtype=torch.bfloat16, device=DEV)
    _warm(s, s.X, Xout, lam, theta, Y=Y)
    F = torch.empty((m, 12), device=DEV)
    N.check(N.lib().xpbd_ws_features(None, N.ptr(s.X), N.ptr(s.XP), N.ptr(s.R), N.ptr(s.ANC), h,
                                     N.ptr(F), m))
    Yf = ws.predict(theta, F).to(torch.bfloat16)     # the pass keeps bfloat16 outputs
    y_k, y_f = Y.float().cpu().numpy(), Yf.float().cpu().numpy()
    assert np.array_equal(y_k, y_f), np.abs(y_k - y_f).max()
    omega = 4.0 / (1.0 + np.exp(-y_k))
    print(f"omega range {omega.min():.2f} .. {omega.max():.2f}")
    assert omega.max() - omega.min() > 0.25
    rows = s.R[:m].cpu().numpy()
    anc = s.ANC[:m].cpu().numpy()
    Fr = ws_ref.features(X, XP, rows, anc, h)
    assert np.allclose(F.cpu().numpy(), Fr, rtol=1e-5, atol=2e-5)
    Xr, lam_r, _ = ws_ref.warm_launch(X, XP, rows, anc, y_k, h, s.jacobi_relax)
    lg = lam.cpu().numpy()
    t = (rows[:, 0].view(np.uint32) >> 28)
    assert (lam_r[t == 2] != 0).sum() > 50                      # violated FLOOR rows pushed
    lerr = np.abs(lg - lam_r).max()
    print(f"lambda0: max |gpu - oracle| {lerr:.3e}, max |lambda0| {np.abs(lam_r).max():.3e}")
    assert lerr <= 1e-4 * np.abs(lam_r).max()
    xg = Xout.cpu().numpy()
    err = np.abs(xg[:, :3] - Xr[:, :3]).max()
    move = np.abs(Xr[:, :3] - X[:, :3]).max()
    print(f"positions: max |gpu - oracle| {err:.3e}, largest move {move:.3e}")
    assert err <= 1e-6 + 1e-4 * move


@pytest.mark.parametrize("predict", [0, 1])
def test_network_pass_in_64_row_groups_equals_features_then_forward(predict):
    """On a row set large enough for 64-row groups (96^2 cloth, ~54 k rows,
    a ragged last group) the network pass gives bit for bit what
    xpbd_ws_features -> xpbd_mlp_forward gives at the same positions, with
    and without integrate folded in."""
    import torch

    from xpbd_amd import _native as N
    from xpbd_amd import warmstart as ws

    s = _cloth(96)
    ws._prepare(s)
    _perturbed_state(s, 3, below_floor=True)
    m, n = s.num_constraints, s.num_bodies
    assert m >= 64 * 256 and m % 64 != 0
    h = float(np.float32(s.dt / s.substeps))
    theta = _theta(11, b3=0.5)
    img = torch.empty(int(N.lib().xpbd_mlp_image_bytes()), dtype=torch.uint8, device=DEV)
    N.check(N.lib().xpbd_mlp_pack(None, N.ptr(theta), N.ptr(img)))
    Y = torch.full((m + 7,), 7.0, dtype=torch.bfloat16, device=DEV)
    N.check(N.lib().xpbd_ws_net(None, N.ptr(s.X), N.ptr(s.XP), N.ptr(s.V), N.ptr(s.R),
                                N.ptr(s.ANC), N.ptr(img), N.ptr(Y), h, *s.gravity, None, m,
                                predict, None, 0))
    X, XP = s.X, s.XP
    if predict:   # features at the integrated positions
        X, XP, V = s.X.clone(), torch.empty_like(s.XP), s.V.clone()
        N.check(N.lib().xpbd_integrate_to(None, N.ptr(s.X), N.ptr(X), N.ptr(V), N.ptr(XP),
                                          *s.gravity, h, None, n))
    F = torch.empty((m, 12), device=DEV)
    N.check(N.lib().xpbd_ws_features(None, N.ptr(X), N.ptr(XP), N.ptr(s.R), N.ptr(s.ANC), h,
                                     N.ptr(F), m))
    Yf = ws.predict(theta, F).to(torch.bfloat16)     # the pass keeps bfloat16 outputs
    assert torch.equal(Y[:m], Yf)
    assert bool((Y[m:] == 7.0).all())            # nothing written past the rows


def test_zero_start_is_the_plain_solver():
    """With the network's output pushed to -inf (omega = 0) the warm-start
    launch leaves the predicted positions as they are, writes zero
    multipliers, and two frames of the learned frame plan equal two frames
    of the plain plan bit for bit (cloth meeting the floor: active FLOOR
    rows with friction)."""
    import torch

    from xpbd_amd import _native as N
    from xpbd_amd import warmstart as ws

    s = _cloth(24)
    ws._prepare(s)
    _perturbed_state(s, 6, below_floor=True)
    m, n = s.num_constraints, s.num_bodies
    h = float(np.float32(s.dt / s.substeps))
    theta = _theta(1, b3=-1e4)
    X0 = s.X.clone()
    Xo = torch.empty_like(s.X)
    lam = torch.ones(m, device=DEV)
    _warm(s, X0, Xo, lam, theta, predict=1)
    Xi = X0.clone()
    V1 = s.V.clone()
    XP1 = torch.empty_like(s.XP)
    N.check(N.lib().xpbd_integrate_to(None, N.ptr(Xi), N.ptr(Xi), N.ptr(V1), N.ptr(XP1),
                                      *s.gravity, h, None, n))
    assert torch.equal(Xo, Xi)
    assert torch.equal(s.XP, XP1)
    assert not bool(lam.any())
    for fuse in (True, False):
        a, b = _cloth(24), _cloth(24)
        for c in (a, b):
            c.fuse_first = fuse
            c.block_frames = False
            c.step(18)                                           # cloth reaches the floor
        a.set_warm_start(theta)
        a.step(2)
        b.step(2)
        assert np.array_equal(a.positions(), b.positions()), fuse
        assert np.array_equal(a.velocities(), b.velocities()), fuse


def test_warm_launch_folded_integrate_equals_integrate_then_warm():
    """The network pass and warm-start launch with integrate folded in equal
    integrate followed by both bit for bit (positions, multipliers, network
    outputs)."""
    import torch

    from xpbd_amd import _native as N
    from xpbd_amd import warmstart as ws

    s = _cloth(24)
    ws._prepare(s)
    _perturbed_state(s, 8, below_floor=True)
    m, n = s.num_constraints, s.num_bodies
    h = float(np.float32(s.dt / s.substeps))
    theta = _theta(3, b3=-0.5)
    X0, V0, XP0 = s.X.clone(), s.V.clone(), s.XP.clone()
    res = []
    for predict in (1, 0):
        s.X.copy_(X0), s.V.copy_(V0), s.XP.copy_(XP0)
        if not predict:
            N.check(N.lib().xpbd_integrate_to(None, N.ptr(s.X), N.ptr(s.X), N.ptr(s.V),
                                              N.ptr(s.XP), *s.gravity, h, None, n))
        Xo = torch.empty_like(s.X)
        lam = torch.zeros(m, device=DEV)
        Y = torch.zeros(m, dtype=torch.bfloat16, device=DEV)
        _warm(s, s.X if not predict else X0.clone(), Xo, lam, theta, Y=Y, predict=predict)
        res.append((Xo, lam, Y, s.XP.clone()))
    for name, a, b in zip(("X", "lambda", "Y", "x_prev"), *res):
        assert torch.equal(a, b), name


@pytest.fixture(scope="module")
def trained():
    from xpbd_amd import warmstart as ws

    log = []
    theta = ws.fit(_cloth, substeps=120, settle_frames=2, steps=800, log=log)
    return theta, log[0]


def test_training_lowers_the_loss_and_data_stay_finite(trained):
    """Data generation (plain, then on-policy) and training on the GPU: the
    final loss is finite, the on-policy round's states stayed finite, and the
    trained network fits fresh data better than its initial parameters."""
    import torch

    from xpbd_amd import warmstart as ws

    theta, log = trained
    print("fit log", log)
    assert log["round2_data_finite"]
    assert bool(torch.isfinite(theta).all())
    assert all(np.isfinite(v) for _, v in log["round1_loss"] + log["round2_loss"])
    F, T = ws.collect(_cloth(), 30, every=3)
    y0 = ws.predict(torch.from_numpy(ws.init_params(0)).to(F.device), F)
    y1 = ws.predict(theta, F)
    e0, e1 = float(((y0 - T) ** 2).mean()), float(((y1 - T) ** 2).mean())
    print(f"held-out MSE: initial {e0:.3f}, trained {e1:.3f}")
    assert e1 < 0.9 * e0


def test_learned_start_reaches_zero_start_residual_with_half_the_iterations(trained):
    """On a fresh run of the same scene, each start on its own trajectory:
    RMS residual after the warm-start launch + k iterations against k and 2k
    iterations from zero."""
    from xpbd_amd import warmstart as ws

    theta, _ = trained
    s = _cloth()
    s.step(20)
    res, fallback = ws.Evaluator(s, theta).run(substeps=60, plain_iters=(1, 2, 4),
                                               learned_iters=(1, 2))
    print("rms residual", res, "fallbacks", fallback)
    assert res["learned_1"] < res["plain_1"]
    assert res["learned_2"] < res["plain_2"]
    assert res["learned_1"] <= res["plain_2"]
    assert res["learned_2"] <= res["plain_4"]
    # the network pass once per frame (15 substeps), its outputs reused by
    # the other substeps with their own first plain steps
    amort, fb = ws.Evaluator(s, theta).run(substeps=60, plain_iters=(), learned_iters=(1, 2),
                                           net_every=15)
    print("network once per frame: rms residual", amort)
    assert amort["learned_1"] <= res["plain_2"]
    assert amort["learned_2"] <= res["plain_4"]


@pytest.mark.parametrize("fuse,every", [(True, 1), (False, 1), (True, 5)])
def test_frame_plan_with_warm_start_matches_eager_launches(trained, fuse, every):
    """Two frames of the captured plan with the warm start (integrate folded
    into the warm-start launch, or not; network pass on every substep or on
    every 5th) equal the eager per-substep launches (integrate, network
    pass, warm-start launch, plain launches) bit for bit."""
    import torch

    from xpbd_amd import warmstart as ws

    theta, _ = trained
    a, b = _cloth(), _cloth()
    for s in (a, b):
        s.fuse_first = fuse
        s.step(3)
    a.warm_start_every = every
    a.set_warm_start(theta)
    a.step(2)
    e = ws._Eager(b)
    th = torch.as_tensor(theta).to(DEV)
    for k in range(2 * b.substeps):
        e.integrate()
        e.iterate(b.iterations, theta=th, net=(k % b.substeps) % every == 0)
    assert np.array_equal(a.positions(), b.positions())
    assert np.array_equal(a.velocities(), b.velocities())


def _stretched_chain():
    """200 free particles (no gravity) 0.3 m apart joined by rows of rest
    0.2 m, rigid and compliant (1e-4) alternating: every row starts far
    from its solution.  Frames of two substeps of 1/900 s."""
    from xpbd_amd import Body, Solver

    s = Solver(dt=2 / 900, substeps=2, iterations=2, solve_mode="jacobi", device=DEV,
               gravity=(0.0, 0.0, 0.0))
    for k in range(200):
        s.add_particle((k * 0.3, 0.0, 0.0), 1.0, collide=False)
    for k in range(199):
        s.add_distance(Body(k), Body(k + 1), rest=0.2, compliance=0.0 if k % 2 else 1e-4)
    return s


def test_judge_keeps_the_trained_start_and_drops_a_poor_one_within_a_frame(trained):
    """The plain zero start as judge in the frame plan.  With the trained
    network the learned start stays on frame after frame.  On the stretched
    chain a hand-built network that starts the rigid rows from zero and
    the compliant rows from 4 times their first step is worse than the zero
    start on some frames: every frame, the start the plan runs next is the
    better one on the frame's first substep, as the eager Evaluator measures
    it from the same state (a poor start is off from the frame's second
    substep on), and it is switched off at least once."""
    from xpbd_amd import warmstart as ws

    theta, _ = trained
    s = _cloth()
    s.warm_start_judge = 1
    s.step(10)
    s.set_warm_start(theta)
    s.step(12)
    st = s.warm_start_stats()
    print("trained start:", st)
    assert st["judged_learned"] + st["judged_zero"] == 12
    assert st["judged_learned"] >= 10 and st["learned_on"]
    # y = -12 + 3 GELU(GELU(10 f6)), f6 = alpha~ / (w + alpha~): about -12
    # (lambda0 = 0) on rigid rows, +17 (lambda0 = 4 x the first step) on
    # the compliant ones
    th = np.zeros(ws.N_PARAMS, np.float32)
    th[6 * 64] = 10.0               # W1[6][0]
    th[12 * 64 + 64] = 1.0          # W2[0][0]
    th[12 * 64 + 64 + 64 * 64 + 64] = 3.0    # W3[0]
    th[-1] = -12.0                  # b3
    import torch

    bad = torch.from_numpy(th).to(DEV)
    c, probe = _stretched_chain(), _stretched_chain()
    c.warm_start_judge = 1
    c.set_warm_start(bad)
    offs = 0
    for f in range(16):
        x, v = c.positions(), c.velocities()
        probe.set_state(x, v)
        res, _ = ws.Evaluator(probe, bad).run(substeps=1, plain_iters=(2,), learned_iters=(2,))
        c.step()
        on = c.warm_start_stats()["learned_on"]
        lr, pr = res["learned_2"], res["plain_2"]
        print(f"frame {f}: learned {lr:.4e} plain {pr:.4e} -> learned on {on}")
        if abs(lr - pr) > 1e-6 * pr:
            assert on == (lr <= pr), f
            offs += not on
    st = c.warm_start_stats()
    print("poor start:", st)
    assert offs >= 1 and st["switches"] >= 1
    assert st["judged_learned"] + st["judged_zero"] == 16
    assert np.all(np.isfinite(c.positions()))


@pytest.mark.parametrize("graph", [True, False])
def test_start_switched_off_runs_the_plain_solver_bit_for_bit(trained, graph):
    """With the flag at 'zero start' (and no judge to turn it back on) the
    warm-start launches leave the multipliers at zero and the positions as
    predicted, so the plan gives, bit for bit, the states and multipliers of
    the plain plan (cloth on the floor, friction active)."""
    a, b = _cloth(), _cloth()
    for s in (a, b):
        s.use_graph = graph
        s.step(30)
    a.warm_start_judge = False
    a.set_warm_start(_theta(3, b3=12.0))
    a._flush()
    a._get_plan()
    a.WSON.zero_()
    for _ in range(5):
        a.step()
        b.step()
    assert np.array_equal(a.positions(), b.positions())
    assert np.array_equal(a.velocities(), b.velocities())
    assert np.array_equal(a.lambdas(), b.lambdas())
    assert a.warm_start_stats()["learned_on"] is False


def test_ws_residual_matches_numpy_oracle():
    """xpbd_ws_residual against the float64 restatement on a cloth half
    below the floor with random multipliers."""
    import torch

    from oracle import ws_ref
    from xpbd_amd import _native as N
    from xpbd_amd import warmstart as ws

    s = _cloth(16)
    ws._prepare(s)
    X, _ = _perturbed_state(s, 3, below_floor=True)
    m = s.num_constraints
    h = float(np.float32(s.dt / s.substeps))
    rng = np.random.default_rng(4)
    lam = rng.normal(0, 1e-3, m).astype(np.float32)
    LAM = torch.from_numpy(lam).to(DEV)
    out = torch.zeros(1, dtype=torch.float64, device=DEV)
    N.check(N.lib().xpbd_ws_residual(None, N.ptr(s.X), N.ptr(s.R), N.ptr(s.ANC), N.ptr(LAM), h,
                                     N.ptr(out), m))
    r_ref = ws_ref.residual_sum(X, s.R[:m].cpu().numpy(), s.ANC[:m].cpu().numpy(), lam, h)
    assert np.isclose(float(out.item()), r_ref, rtol=1e-4)


@pytest.mark.parametrize("fuse", [True, False])
def test_warm_start_with_device_contact_rows_matches_eager(trained, fuse):
    """Particle scenes of a small batched world (cloths, chains, sphere
    piles whose CONTACT rows are generated on the device each frame, gravity
    scaled per scene): two frames of the captured plan with the learned
    start (network once per frame; integrate folded in or not) agree with
    the eager per-substep launches (contacts per frame, integrate, network
    pass, warm-start launch, device-row projection before every iteration)
    to float-atomic rounding, and the piles stay above the floor."""
    import torch

    from xpbd_amd import warmstart as ws
    from xpbd_amd.batch import BatchedWorld, draw

    theta, _ = trained
    spec = draw(12, 3, mix={"cloth": 0.4, "pile": 0.4, "chain": 0.2})
    assert any(sc.kind == "pile" for sc in spec.scenes)
    a = BatchedWorld(spec, device=DEV).particles
    b = BatchedWorld(spec, device=DEV).particles
    for s in (a, b):
        s.fuse_first = fuse
        s.block_frames = False
        s.step(3)
    assert a.CROWS is not None and a.n_contacts > 0
    a.warm_start_every = 15
    a.set_warm_start(theta)
    a.step(2)
    e = ws._Eager(b)
    th = torch.as_tensor(theta).to(DEV)
    for k in range(2 * b.substeps):
        if k % b.substeps == 0:
            e.contacts()
        e.integrate()
        e.iterate(b.iterations, theta=th, net=(k % b.substeps) == 0)
    xa, xb = a.positions(), b.positions()
    assert np.all(np.isfinite(xa))
    assert np.abs(xa - xb).max() <= 1e-4
    assert np.abs(a.velocities() - b.velocities()).max() <= 5e-3
    piles = [sc for sc in spec.scenes if sc.kind == "pile"]
    for sc in piles:
        assert xa[sc.first:sc.first + sc.count, 1].min() > 0.05 - 0.01


@pytest.mark.parametrize("off", [False, True])
def test_one_launch_learned_substep_equals_start_launch_then_iteration(trained, off):
    """One-iteration substeps with the learned start on a small cloth: the
    plan's single launch per substep (xpbd_jacobi_tile_pair_warm: the start
    on the tile and its ring, then the iteration) gives, bit for bit, the
    states and multipliers of the plan with a start launch and an iteration
    launch per substep; the judge runs every 2nd frame, the cloth rests on
    the floor (friction active).  off: the start flag held at 'zero start'."""
    theta, _ = trained
    a, b = _cloth(), _cloth()
    for s in (a, b):
        s.step(30)
        s.iterations = 1
        s.warm_start_every = s.substeps
        s.warm_start_judge = False if off else 2
        s.set_warm_start(theta)
    a.tile_pair = "auto"
    b.tile_pair = False
    for s in (a, b):
        s._flush()
        s._get_plan()
        if off:
            s.WSON.zero_()
    assert a._plan_pair and not b._plan_pair
    for f in range(6):
        a.step()
        b.step()
        assert np.array_equal(a.positions(), b.positions()), f
        assert np.array_equal(a.velocities(), b.velocities()), f
        assert np.array_equal(a.lambdas(), b.lambdas()), f
    assert a.warm_start_stats() == b.warm_start_stats()
