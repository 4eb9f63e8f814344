"""Learned lambda warm start for the 3-DOF Jacobi solver (no reference
counterpart; it replaces the per-substep zero reset of the multipliers,
src/xpbd3d/solver.py:431).

* ``collect``  runs the plain solver and, at sampled substeps, records the
  12 row features (kernel ``xpbd_ws_features``) and the target
  y* = slog(lambda* (w + alpha~) / 1 mm), lambda* = the multipliers after
  ``target_iters`` plain Jacobi iterations of that substep from lambda = 0
  (run on a copy of the state, so the trajectory itself is the plain one).
* ``train``    fits the 12-64-64-1 MLP on the GPU with the library's own BF16
  MFMA forward/backward (``xpbd_mlp_grad``) and Adam (``xpbd_mlp_adam``).
* ``Evaluator`` steps one substep at a time with the plain zero start or the
  learned start (features -> MLP -> lambda0 and the matching position
  offset -> iterations continuing from lambda0) and measures the RMS
  constraint residual after the last iteration; the plain solver is the
  judge, and the fallback keeps the plain result whenever the learned start
  ends with a larger residual.
"""
from __future__ import annotations

import numpy as np

from . import _native as N

LEN = 1e-3
N_PARAMS = 5057


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
        s = self.s
        N.check(self.lib.xpbd_ws_features(N.current_stream(), N.ptr(s.X), N.ptr(s.V), N.ptr(s.R),
                                          N.ptr(s.ANC), N.ptr(s.LAM), self.h, N.ptr(F), self.m))

    def warm(self, theta, F, Y):
        s = self.s
        self.features(F)
        N.check(self.lib.xpbd_mlp_forward(N.current_stream(), N.ptr(theta), N.ptr(F), N.ptr(Y),
                                          self.m))
        N.check(self.lib.xpbd_ws_apply(N.current_stream(), N.ptr(s.X), N.ptr(s.V), N.ptr(s.R),
                                       N.ptr(s.ANC), N.ptr(Y), self.h, N.ptr(s.LAM), N.ptr(s.ACC),
                                       self.m))
        N.check(self.lib.xpbd_particles_pass(N.current_stream(), N.ptr(s.X), N.ptr(s.V),
                                             N.ptr(s.XP), N.ptr(s.ACC), 16, s.jacobi_relax,
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


def collect(s, substeps: int, every: int = 1, target_iters: int = 64, iters: int | None = None):
    """Run the plain solver for ``substeps`` substeps; every ``every``-th
    substep append (features, target) for all rows.  Returns (F, T) device
    tensors of shape (k m, 12) and (k m,)."""
    import torch

    _prepare(s)
    e = _Eager(s)
    iters = s.iterations if iters is None else iters
    scale = row_scale(s)
    F_all, T_all = [], []
    F = torch.empty((e.m, 12), device=s.device)
    for k in range(substeps):
        e.integrate()
        if k % every == 0:
            e.features(F)
            Xc = s.X.clone()
            Lc = torch.zeros_like(s.LAM)
            Ac = torch.zeros_like(s.ACC)
            e.iterate(Xc, Lc, Ac, target_iters, first=True)
            F_all.append(F.clone())
            T_all.append(_slog(Lc[:e.m] * scale / LEN))
        e.iterate(s.X, s.LAM, s.ACC, iters, first=True)
        e.finalize()
    return torch.cat(F_all), torch.cat(T_all)


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
    """Plain vs learned start on the same trajectory: at each substep both
    start from the plain run's state; residuals after k iterations are
    compared, and the fallback keeps the plain result when the learned one is
    worse."""

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
        res = {("plain", k): [] for k in plain_iters}
        res.update({("learned", k): [] for k in learned_iters})
        fallback = {k: 0 for k in learned_iters}
        for _ in range(substeps):
            e.integrate()
            x0, v0, lam_prev = s.X.clone(), s.V.clone(), s.LAM.clone()
            plain_k = {}
            for k in plain_iters:
                s.X.copy_(x0)
                s.V.copy_(v0)
                s.ACC.zero_()
                e.iterate(s.X, s.LAM, s.ACC, k, first=True)
                plain_k[k] = e.residual()
                res[("plain", k)].append(plain_k[k])
            for k in learned_iters:
                s.X.copy_(x0)
                s.V.copy_(v0)
                s.LAM.copy_(lam_prev)
                s.ACC.zero_()
                e.warm(self.theta, self.F, self.Y)
                e.iterate(s.X, s.LAM, s.ACC, k, first=False)
                r = e.residual()
                res[("learned", k)].append(r)
                if k in plain_k and r > plain_k[k]:
                    fallback[k] += 1
            # continue the trajectory with the reference's plain 2-iteration substep
            s.X.copy_(x0)
            s.V.copy_(v0)
            s.ACC.zero_()
            e.iterate(s.X, s.LAM, s.ACC, max(plain_iters[0], s.iterations), first=True)
            e.finalize()
        return ({f"{a}_{k}": float(np.sqrt(np.mean(np.square(v)))) for (a, k), v in res.items()},
                fallback)
