"""Learned lambda warm start for the 3-DOF Jacobi solver (no reference
counterpart; it replaces the per-substep zero reset of the multipliers,
src/xpbd3d/solver.py:431).

* ``collect``  runs the solver and, at sampled substeps, records the row
  features (kernel ``xpbd_ws_features``) and the target y* =
  logit(omega / 8), omega = beta lambda* / lambda1: lambda* = the
  multipliers after ``target_iters`` plain Jacobi iterations of that substep
  from lambda = 0 (run on a copy of the state), beta = 0.5, lambda1 = the
  row's first plain step -C / (w + alpha~); the deployed start is
  lambda0 = 8 sigmoid(y) lambda1, bounded by the row's current violation;
* ``fit``      two rounds of collect + train: on the plain trajectory, then on
  the learned start's own trajectory;
* ``train``    fits the 12-64-64-1 MLP on the GPU with the library's own BF16
  MFMA forward/backward (``xpbd_mlp_grad``) and Adam (``xpbd_mlp_adam``);
* ``Evaluator`` runs the plain zero start and the learned start (features ->
  MLP -> lambda0 and the matching position offset -> iterations continuing
  from lambda0) on their own trajectories and measures the RMS constraint
  residual after each substep's iterations;
* ``Guard``    keeps the plain solver as fallback and judge in deployment.

The previous-substep lambda feature slot is left at zero: measured on
cloth, carrying it in feeds every warm-started substep's multipliers into
the next prediction and the scene blows up within a few frames.
"""
from __future__ import annotations

import numpy as np

from . import _native as N

LEN = 1e-3
N_PARAMS = 5057
OMEGA_MAX = 8.0         # kOmegaMax of csrc/warmstart.hip


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal weights, zero biases; layout W1[12][64] b1 W2[64][64] b2 W3[64] b3."""
    rng = np.random.default_rng(seed)
    th = np.zeros(N_PARAMS, np.float32)
    th[0:768] = rng.normal(0, np.sqrt(2.0 / 12), 768)
    th[832:832 + 4096] = rng.normal(0, np.sqrt(2.0 / 64), 4096)
    th[4992:5056] = rng.normal(0, np.sqrt(1.0 / 64), 64)
    return th


def _slog(x):
    import torch

    return torch.sign(x) * torch.log1p(torch.abs(x))


def _slog_inv(y):
    import torch

    return torch.sign(y) * torch.expm1(torch.abs(y))


def target_logit(F, lam_star, scale, beta: float):
    """Network target y* = logit(omega / OMEGA_MAX), omega = beta lambda* /
    lambda1 clipped to (0, OMEGA_MAX), lambda1 = -C / (w + alpha~) the first
    plain step from zero (C recovered from feature 4); rows without a step
    (separated one-sided rows, w + alpha~ = 0) get y* = 0."""
    import torch

    C = _slog_inv(F[:, 4]) * LEN
    one_sided = (F[:, 2] > 0.5) | (F[:, 3] > 0.5)
    lam1 = torch.where(scale > 0, -C / torch.clamp(scale, min=1e-30), torch.zeros_like(C))
    lam1 = torch.where(one_sided & (C >= 0), torch.zeros_like(lam1), lam1)
    valid = lam1.abs() > 1e-12
    ratio = torch.where(valid, beta * lam_star / torch.where(valid, lam1, torch.ones_like(lam1)),
                        torch.zeros_like(lam1))
    om = torch.clamp(ratio, 0.01, OMEGA_MAX - 0.01)
    y = torch.log(om / (OMEGA_MAX - om))
    return torch.where(valid, y, torch.zeros_like(y))


def row_scale(s) -> "torch.Tensor":
    """w + alpha~ per device row of a flushed Jacobi ``Solver``."""
    import torch

    R = s.R[:s.num_constraints]
    a = (R[:, 0] & 0x0FFFFFFF).long()
    b = R[:, 1].long()
    t = (R[:, 0] >> 28) & 0xF
    w = s.X[a, 3] + torch.where(b >= 0, s.X[b.clamp(min=0), 3], torch.zeros_like(s.X[a, 3]))
    w = torch.where((t == 1) | (t == 2), s.X[a, 3], w)            # ATTACH / FLOOR: one body
    h = float(np.float32(s.dt / s.substeps))
    comp = R[:, 3].view(torch.float32)
    alpha = torch.where((t == 0) | (t == 1), comp / (h * h), torch.zeros_like(comp))
    return w + alpha


class _Eager:
    """Per-substep launches of the window-Jacobi path through the C-ABI."""

    def __init__(self, s):
        s._flush()
        self.s = s
        self.lib = N.lib()
        self.h = float(np.float32(s.dt / s.substeps))
        self.m = s.num_constraints
        self.n = s.num_bodies

    def integrate(self, X=None):
        s = self.s
        N.check(self.lib.xpbd_particles_pass(N.current_stream(), N.ptr(s.X if X is None else X),
                                             N.ptr(s.V), N.ptr(s.XP), N.ptr(s.ACC), 8,
                                             s.jacobi_relax, self.h, s.damping, s.max_speed,
                                             *s.gravity, self.n))

    def iterate(self, X, LAM, ACC, iters, first):
        s = self.s
        rpt, win = s._tiling
        for it in range(iters):
            N.check(self.lib.xpbd_solve_jacobi(N.current_stream(), N.ptr(X), N.ptr(s.XP),
                                               N.ptr(s.R), N.ptr(s.ANC), N.ptr(LAM), N.ptr(ACC),
                                               self.h, int(first and it == 0), self.m, rpt, win))
            N.check(self.lib.xpbd_particles_pass(N.current_stream(), N.ptr(X), N.ptr(s.V),
                                                 N.ptr(s.XP), N.ptr(ACC), 1, s.jacobi_relax,
                                                 self.h, s.damping, s.max_speed, *s.gravity,
                                                 self.n))

    def finalize(self):
        s = self.s
        N.check(self.lib.xpbd_particles_pass(N.current_stream(), N.ptr(s.X), N.ptr(s.V),
                                             N.ptr(s.XP), N.ptr(s.ACC), 2 | 4, s.jacobi_relax,
                                             self.h, s.damping, s.max_speed, *s.gravity, self.n))

    def features(self, F):
        """Row features; the previous-substep lambda slot stays zero (fed
        back from warm-started substeps it drives the predictions upward
        from substep to substep until the scene blows up)."""
        s = self.s
        N.check(self.lib.xpbd_ws_features(N.current_stream(), N.ptr(s.X), N.ptr(s.V), N.ptr(s.R),
                                          N.ptr(s.ANC), None, self.h, N.ptr(F), self.m))

    def warm(self, theta, F, Y):
        s = self.s
        self.features(F)
        N.check(self.lib.xpbd_mlp_forward(N.current_stream(), N.ptr(theta), N.ptr(F), N.ptr(Y),
                                          self.m))
        N.check(self.lib.xpbd_ws_apply(N.current_stream(), N.ptr(s.X), N.ptr(s.V), N.ptr(s.R),
                                       N.ptr(s.ANC), N.ptr(Y), self.h, N.ptr(s.LAM), N.ptr(s.ACC),
                                       self.m))
        N.check(self.lib.xpbd_particles_pass(N.current_stream(), N.ptr(s.X), N.ptr(s.V),
                                             N.ptr(s.XP), N.ptr(s.ACC), 1, s.jacobi_relax,
                                             self.h, s.damping, s.max_speed, *s.gravity, self.n))

    def residual(self, X=None, LAM=None) -> float:
        import torch

        s = self.s
        out = torch.zeros(1, dtype=torch.float64, device=s.device)
        N.check(self.lib.xpbd_ws_residual(N.current_stream(), N.ptr(s.X if X is None else X),
                                          N.ptr(s.R), N.ptr(s.ANC),
                                          N.ptr(s.LAM if LAM is None else LAM), self.h,
                                          N.ptr(out), self.m))
        return float(np.sqrt(out.item() / max(self.m, 1)))


def _prepare(s):
    if s.solve_mode != "jacobi":
        raise ValueError("the learned warm start runs with solve_mode='jacobi'")
    s.jacobi_kernel = "window"
    s._dirty = True
    s._flush()


TARGET_SCALE = 0.5


def collect(s, substeps: int, every: int = 1, target_iters: int = 64, iters: int | None = None,
            theta=None, target_scale: float = TARGET_SCALE):
    """Run the solver for ``substeps`` substeps; every ``every``-th substep
    append (features, target) for all rows.  Returns (F, T) device tensors of
    shape (k m, 12) and (k m,).

    Target: ``target_logit`` of beta lambda*, lambda* = the multipliers after
    ``target_iters`` plain iterations from zero at that substep (run on a
    copy of the state), beta = ``target_scale``: the network predicts the
    multiple omega of the row's first plain step that the start should take.
    A full lambda* start overshoots (measured on cloth: reusing the previous
    substep's lambda explodes within a few frames, half of it matches the
    plain solver with twice the iterations), so the target is damped.

    ``theta`` None: the trajectory is the plain solver's; ``theta`` given: the
    trajectory runs the learned start (an on-policy round, whose states are
    the ones the deployed network meets)."""
    import torch

    _prepare(s)
    e = _Eager(s)
    iters = s.iterations if iters is None else iters
    scale = row_scale(s)
    F_all, T_all = [], []
    F = torch.empty((e.m, 12), device=s.device)
    Y = torch.empty(e.m, device=s.device)
    for k in range(substeps):
        e.integrate()
        if k % every == 0:
            e.features(F)
            Xc = s.X.clone()
            Lc = torch.zeros_like(s.LAM)
            Ac = torch.zeros_like(s.ACC)
            e.iterate(Xc, Lc, Ac, target_iters, first=True)
            F_all.append(F.clone())
            T_all.append(target_logit(F, Lc[:e.m], scale, float(target_scale)))
        if theta is None:
            e.iterate(s.X, s.LAM, s.ACC, iters, first=True)
        else:
            e.warm(theta, F, Y)
            e.iterate(s.X, s.LAM, s.ACC, iters, first=False)
        e.finalize()
    return torch.cat(F_all), torch.cat(T_all)


def fit(make_scene, substeps: int = 120, settle_frames: int = 2, steps: int = 800,
        batch: int = 16384, seed: int = 0, log: list | None = None):
    """Two rounds: data from the plain solver's trajectory, training, data
    from the learned start's own trajectory, training on both.
    ``make_scene()`` returns a fresh Jacobi ``Solver``; returns theta."""
    import torch

    def fresh():
        s = make_scene()
        s.step(settle_frames)
        return s

    F, T = collect(fresh(), substeps)
    theta, h1 = train(F, T, steps=steps, batch=batch, lr=3e-3, seed=seed, log_every=steps)
    F2, T2 = collect(fresh(), substeps, theta=theta)
    ok = bool(torch.isfinite(F2).all() and torch.isfinite(T2).all())
    if ok:
        theta, h2 = train(torch.cat([F, F2]), torch.cat([T, T2]), steps=steps, batch=batch,
                          lr=1e-3, seed=seed + 1, theta=theta, log_every=steps)
    else:
        h2 = []
    if log is not None:
        log.append({"round1_loss": h1, "round2_loss": h2, "round2_data_finite": ok,
                    "samples": int(F.shape[0] + (F2.shape[0] if ok else 0))})
    return theta


def train(F, T, steps: int = 2000, batch: int = 65536, lr: float = 3e-3, seed: int = 0,
          theta=None, nblocks: int = 256, log_every: int = 0):
    """Adam on minibatches drawn from (F, T); returns (theta, loss history)."""
    import torch

    lib = N.lib()
    dev = F.device
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    th = (torch.from_numpy(init_params(seed)).to(dev) if theta is None else theta.clone())
    mom = torch.zeros_like(th)
    vel = torch.zeros_like(th)
    grad = torch.zeros_like(th)
    loss = torch.zeros(1, device=dev)
    ws = torch.empty(int(lib.xpbd_mlp_train_workspace_bytes(nblocks)), dtype=torch.uint8,
                     device=dev)
    hist = []
    n = F.shape[0]
    b = min(batch, n)
    for k in range(1, steps + 1):
        idx = torch.randint(0, n, (b,), device=dev, generator=g)
        Fb = F[idx].contiguous()
        Tb = T[idx].contiguous()
        N.check(lib.xpbd_mlp_grad(N.current_stream(), N.ptr(th), N.ptr(Fb), N.ptr(Tb), b, nblocks,
                                  N.ptr(ws), N.ptr(grad), N.ptr(loss)))
        N.check(lib.xpbd_mlp_adam(N.current_stream(), N.ptr(th), N.ptr(mom), N.ptr(vel),
                                  N.ptr(grad), lr, 0.9, 0.999, 1e-8, k))
        if log_every and (k % log_every == 0 or k == steps):
            hist.append((k, float(loss.item())))
    return th, hist


def predict(theta, F):
    import torch

    Y = torch.empty(F.shape[0], device=F.device)
    N.check(N.lib().xpbd_mlp_forward(N.current_stream(), N.ptr(theta), N.ptr(F), N.ptr(Y),
                                     F.shape[0]))
    return Y


class Evaluator:
    """Plain zero start against the learned start, each on its own trajectory
    from the same initial state (what a deployed solver sees, including its
    own previous-substep lambda).  Per substep the RMS residual after the
    iterations is recorded; ``run`` returns the RMS over the substeps per
    configuration and, per learned iteration count k, the number of substeps
    whose residual exceeded the plain run's at the same k and substep (the
    substeps on which the fallback would take over).  The state is restored
    afterwards."""

    def __init__(self, s, theta):
        import torch

        _prepare(s)
        self.s = s
        self.e = _Eager(s)
        self.theta = theta
        self.F = torch.empty((self.e.m, 12), device=s.device)
        self.Y = torch.empty(self.e.m, device=s.device)

    def run(self, substeps: int, plain_iters=(1, 2, 4), learned_iters=(1, 2)):
        s, e = self.s, self.e
        keep = [t.clone() for t in (s.X, s.V, s.XP, s.LAM)]

        def restore():
            for dst, src in zip((s.X, s.V, s.XP, s.LAM), keep):
                dst.copy_(src)
            s.ACC.zero_()

        runs = {}
        for kind, its in [("plain", k) for k in plain_iters] + [("learned", k)
                                                                 for k in learned_iters]:
            restore()
            r = []
            for _ in range(substeps):
                e.integrate()
                if kind == "learned":
                    e.warm(self.theta, self.F, self.Y)
                    e.iterate(s.X, s.LAM, s.ACC, its, first=False)
                else:
                    e.iterate(s.X, s.LAM, s.ACC, its, first=True)
                r.append(e.residual())
                e.finalize()
            runs[(kind, its)] = np.asarray(r)
        restore()
        out = {f"{kind}_{k}": float(np.sqrt(np.mean(np.square(v)))) for (kind, k), v in runs.items()}
        fallback = {k: int(np.sum(runs[("learned", k)] > runs[("plain", k)]))
                    for k in learned_iters if ("plain", k) in runs}
        return out, fallback


class Guard:
    """The learned start with the plain zero start as fallback and judge.

    Every ``check_every`` frames one substep is replayed both ways from the
    current state (``iters`` Jacobi iterations each, eager launches, state
    restored afterwards).  The learned start stays on while its RMS residual
    is no larger than the plain one; otherwise the solver runs the zero start
    until the next check.  ``history`` keeps (frame, plain, learned, on)."""

    def __init__(self, s, theta, check_every: int = 30, iters: int | None = None):
        import torch

        _prepare(s)
        self.s = s
        self.theta = torch.as_tensor(theta, dtype=torch.float32).to(s.device).contiguous()
        self.every = max(1, int(check_every))
        self.iters = s.iterations if iters is None else int(iters)
        self.frame = 0
        self.on = False
        self.history: list = []

    def compare(self) -> tuple[float, float]:
        """(plain, learned) RMS residual of one substep from the current state."""
        import torch

        s = self.s
        e = _Eager(s)
        keep = [t.clone() for t in (s.X, s.V, s.XP, s.LAM)]
        F = torch.empty((e.m, 12), device=s.device)
        Y = torch.empty(e.m, device=s.device)

        def restore():
            for dst, src in zip((s.X, s.V, s.XP, s.LAM), keep):
                dst.copy_(src)
            s.ACC.zero_()

        out = []
        for learned in (False, True):
            restore()
            e.integrate()
            if learned:
                e.warm(self.theta, F, Y)
            e.iterate(s.X, s.LAM, s.ACC, self.iters, first=not learned)
            out.append(e.residual())
        restore()
        return out[0], out[1]

    def step(self, frames: int = 1) -> None:
        for _ in range(frames):
            if self.frame % self.every == 0:
                rp, rl = self.compare()
                on = rl <= rp
                if on != self.on:
                    self.s.set_warm_start(self.theta if on else None)
                    self.on = on
                self.history.append((self.frame, rp, rl, on))
            self.s.step()
            self.frame += 1
