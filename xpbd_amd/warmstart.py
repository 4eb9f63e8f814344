"""Learned lambda warm start for the 3-DOF Jacobi solver (no reference
counterpart; it replaces the per-substep zero reset of the multipliers,
src/xpbd3d/solver.py:431).

Each substep starts with the network pass (``xpbd_ws_net``: features and
MLP output y per row) and the warm-start launch (``xpbd_jacobi_tile_warm``):
every row starts from lambda0 = 4 sigmoid(y) lambda1, lambda1 = -C / (w +
alpha~) the row's first plain step (one-sided rows only while violated),
and every particle moves by the mean over its rows of w grad C lambda0 --
the position the Jacobi sweep reaches when its multipliers add up to
lambda0.  The substep's Jacobi iterations then continue from lambda0
(first = 0).  With k iterations a learned substep is k + 2 launches.

* ``collect``  runs the solver and, at sampled substeps, records the row
  features (kernel ``xpbd_ws_features``, the same features the launch
  computes) and the target y* = logit(omega / 4), omega = beta lambda* /
  lambda1 clipped to [0.2, 3.8]: lambda* = the multipliers after
  ``target_iters`` plain iterations of that substep from zero (run on a copy
  of the state), beta = 0.5 (a damped target: the full lambda* overshoots);
* ``fit``      two rounds of collect + train: on the plain trajectory, then on
  the learned start's own trajectory;
* ``train``    fits the 12-64-64-1 MLP on the GPU with the library's own BF16
  MFMA forward/backward (``xpbd_mlp_grad``) and Adam (``xpbd_mlp_adam``);
* ``Evaluator`` runs the plain zero start and the learned start on their
  own trajectories and measures the RMS constraint residual after each
  substep's iterations;

Measured on cloth: the previous-substep lambda feature (slot 7, left at
zero) feeds every warm-started substep's multipliers into the next
prediction and the scene blows up within a few frames; folding the start
into the first iteration's update instead of a launch of its own (a learned
per-row over-relaxation) lost to the zero start at one iteration.
"""
from __future__ import annotations

import numpy as np

from . import _native as N

LEN = 1e-3
N_PARAMS = 5057
OMEGA_MAX = 4.0         # kOmegaMax of csrc/ws_dev.h
TARGET_CLIP = 0.05      # targets omega / OMEGA_MAX clipped to [0.05, 0.95]


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
    lambda1 clipped to OMEGA_MAX x [TARGET_CLIP, 1 - TARGET_CLIP] (targets
    stay within +-3 and the many rows whose ratio lies beyond the cap do not
    dominate the squared error), lambda1 = -C / (w + alpha~) the first
    plain step from zero (C recovered from feature 4, w + alpha~ = ``scale``);
    rows without a step (separated one-sided rows, w + alpha~ = 0) get
    y* = 0."""
    import torch

    C = _slog_inv(F[:, 4]) * LEN
    one_sided = (F[:, 2] > 0.5) | (F[:, 3] > 0.5)
    lam1 = torch.where(scale > 0, -C / torch.clamp(scale, min=1e-30), torch.zeros_like(C))
    lam1 = torch.where(one_sided & (C >= 0), torch.zeros_like(lam1), lam1)
    valid = lam1.abs() > 1e-12
    ratio = torch.where(valid, beta * lam_star / torch.where(valid, lam1, torch.ones_like(lam1)),
                        torch.zeros_like(lam1))
    om = torch.clamp(ratio, TARGET_CLIP * OMEGA_MAX, (1.0 - TARGET_CLIP) * OMEGA_MAX)
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
    h = np.float32(s.dt / s.substeps)
    inv_dt2 = float(np.float32(1.0) / (h * h))       # float32 arithmetic, as the kernels
    comp = R[:, 3].view(torch.float32)
    alpha = torch.where((t == 0) | (t == 1), comp * inv_dt2, torch.zeros_like(comp))
    return w + alpha


class _Eager:
    """Per-substep tile launches through the C-ABI.  Device-generated
    contact rows (``enable_device_contacts``) are projected before every
    tile launch from a zero start, as in the frame plan; ``contacts()``
    regenerates them (the frame plan does so once per frame)."""

    def __init__(self, s):
        _prepare(s)
        self.s = s
        self.lib = N.lib()
        self.h = float(np.float32(s.dt / s.substeps))
        self.m = s.num_constraints
        self.n = s.num_bodies
        self.tables = tuple(N.ptr(t) for t in s._tile_dev)

    def image(self, theta):
        """The network ``theta`` packed for the network pass (xpbd_mlp_pack)."""
        import torch

        if getattr(self, "_img", None) is None:
            self._img = torch.empty(int(self.lib.xpbd_mlp_image_bytes()), dtype=torch.uint8,
                                    device=self.s.device)
        N.check(self.lib.xpbd_mlp_pack(N.current_stream(), N.ptr(theta), N.ptr(self._img)))
        return self._img

    def contacts(self):
        """This frame's device contact rows (no-op without them)."""
        if self.s.CROWS is not None:
            self.s._device_contact_pass()

    def integrate(self):
        s = self.s
        N.check(self.lib.xpbd_integrate_to(N.current_stream(), N.ptr(s.X), N.ptr(s.X), N.ptr(s.V),
                                           N.ptr(s.XP), *s.gravity, self.h,
                                           N.ptr(s.GS) if s.GS is not None else None, self.n))

    def iterate(self, iters, theta=None, finalize=True, X=None, X2=None, L=None, L2=None,
                Y=None, net=True):
        """``iters`` tile launches from X (default the solver's state),
        ping-ponging with X2; the result is left in X and L.  ``theta``: the
        network pass (output into Y, default the solver's buffer; skipped
        with ``net`` False, reusing Y) and the warm-start launch first, then
        the iterations continue from its multipliers."""
        s = self.s
        X = s.X if X is None else X
        X2 = s.X_B if X2 is None else X2
        L = s.LAM if L is None else L
        L2 = s.LAM_B if L2 is None else L2
        src, dst, ls, ld = X, X2, L, L2
        st = N.current_stream()
        trs, hs, hr = self.tables
        if theta is not None:
            Y = s.WSY if Y is None else Y
            if net:
                img = self.image(theta)
                N.check(self.lib.xpbd_ws_net(st, N.ptr(src), N.ptr(s.XP), N.ptr(s.V),
                                             N.ptr(s.R), N.ptr(s.ANC), N.ptr(img), N.ptr(Y),
                                             self.h, 0.0, 0.0, 0.0, None, self.m, 0, None, 0))
            N.check(self.lib.xpbd_jacobi_tile_warm(
                st, N.ptr(src), N.ptr(dst), N.ptr(s.XP), N.ptr(s.V), N.ptr(s.R), N.ptr(s.ANC),
                N.ptr(ld), trs, hs, hr, N.ptr(Y), self.h, 0.0, 0.0, 0.0, None, s._tile, self.n,
                s.jacobi_relax, 0, None, 0))
            src, dst, ls, ld = dst, src, ld, ls
        anc = N.ptr(s.ANC) if s._has_attach else None   # the variant without anchor loads
        dyn = s.CROWS is not None
        acc = N.ptr(s.ACC) if dyn else None
        for it in range(iters):
            fin = int(finalize and it + 1 == iters)
            if dyn:
                N.check(self.lib.xpbd_solve_jacobi_dyn(
                    st, N.ptr(src), N.ptr(s.XP), N.ptr(s.CROWS), N.ptr(s.CLAM), N.ptr(s.ACC),
                    self.h, int(it == 0), N.ptr(s.CCOUNT), s._ccap))
            N.check(self.lib.xpbd_jacobi_tile(
                st, N.ptr(src), N.ptr(dst), N.ptr(s.XP), N.ptr(s.V), N.ptr(s.R), anc,
                N.ptr(ls), N.ptr(ld), trs, hs, hr, self.h, int(it == 0 and theta is None),
                s._tile, self.n, s.jacobi_relax, fin, s.damping, s.max_speed, acc))
            src, dst, ls, ld = dst, src, ld, ls
        if src is not X:
            X.copy_(src)
            L.copy_(ls)

    def features(self, F):
        """Row features at the integrated positions (velocity (x - x_prev)/h)."""
        s = self.s
        N.check(self.lib.xpbd_ws_features(N.current_stream(), N.ptr(s.X), N.ptr(s.XP), N.ptr(s.R),
                                          N.ptr(s.ANC), self.h, N.ptr(F), self.m))

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
    if s.jacobi_kernel != "tile" or s._dirty:
        s.jacobi_kernel = "tile"
        s._dirty = True
    s._flush()
    if s._tile_dev is None:
        raise ValueError("the learned warm start needs the Jacobi tile kernel")


TARGET_SCALE = 0.5


def collect(s, substeps: int, every: int = 1, target_iters: int = 64, iters: int | None = None,
            theta=None, target_scale: float = TARGET_SCALE, net_every: int = 1,
            pair_with_net: bool = False):
    """Run the solver for ``substeps`` substeps; every ``every``-th substep
    append (features, target) for all rows.  Returns (F, T) device tensors
    of shape (k m, 12) and (k m,).

    ``theta`` None: the trajectory is the plain solver's; ``theta`` given: the
    trajectory runs the learned start (an on-policy round, whose states are
    the ones the deployed network meets), with the network pass on every
    ``net_every``-th substep of each frame as deployed.  ``pair_with_net``:
    each target is paired with the features of the last substep on which
    the network runs (the outputs the deployed start reuses until its next
    network pass) instead of its own substep's."""
    import torch

    e = _Eager(s)
    iters = s.iterations if iters is None else iters
    scale = row_scale(s)
    F_all, T_all = [], []
    F = torch.empty((e.m, 12), device=s.device)
    Fnet = torch.empty_like(F)
    th = None if theta is None else torch.as_tensor(theta, dtype=torch.float32).to(s.device)
    for k in range(substeps):
        if k % s.substeps == 0:
            e.contacts()
        e.integrate()
        net = (k % s.substeps) % max(1, net_every) == 0
        if pair_with_net and net:
            e.features(Fnet)
        if k % every == 0:
            e.features(F)
            Xc, Xc2 = s.X.clone(), s.X.clone()
            Lc, Lc2 = torch.zeros_like(s.LAM), torch.zeros_like(s.LAM)
            e.iterate(target_iters, finalize=False, X=Xc, X2=Xc2, L=Lc, L2=Lc2)
            F_all.append((Fnet if pair_with_net else F).clone())
            T_all.append(target_logit(F, Lc[:e.m], scale, float(target_scale)))
        e.iterate(iters, theta=th, net=net)
    return torch.cat(F_all), torch.cat(T_all)


def fit(make_scene, substeps: int = 120, settle_frames: int = 2, steps: int = 800,
        batch: int = 16384, seed: int = 0, log: list | None = None,
        onpolicy_iters=(1, 2), net_every: int = 1, pair_with_net: bool = False,
        every: int = 1):
    """Two rounds: data from the plain solver's trajectory, training, data
    from the learned start's own trajectories (one per iteration count in
    ``onpolicy_iters``: the states a start followed by one iteration meets
    differ from those after two), training on all of it.  ``net_every`` and
    ``pair_with_net`` (see ``collect``) train a network for a start whose
    network pass runs on every ``net_every``-th substep only; ``every``:
    samples taken on every ``every``-th substep (large scenes).
    ``make_scene()`` returns a fresh Jacobi ``Solver``; returns theta."""
    import torch

    def fresh():
        s = make_scene()
        s.step(settle_frames)
        return s

    F, T = collect(fresh(), substeps, every=every, net_every=net_every,
                   pair_with_net=pair_with_net)
    theta, h1 = train(F, T, steps=steps, batch=batch, lr=3e-3, seed=seed, log_every=steps)
    Fs, Ts = [F], [T]
    ok = True
    for k in onpolicy_iters:
        F2, T2 = collect(fresh(), substeps, every=every, theta=theta, iters=k,
                         net_every=net_every, pair_with_net=pair_with_net)
        if bool(torch.isfinite(F2).all() and torch.isfinite(T2).all()):
            Fs.append(F2)
            Ts.append(T2)
        else:
            ok = False
    h2 = []
    if len(Fs) > 1:
        theta, h2 = train(torch.cat(Fs), torch.cat(Ts), steps=steps, batch=batch, lr=1e-3,
                          seed=seed + 1, theta=theta, log_every=steps)
    if log is not None:
        log.append({"round1_loss": h1, "round2_loss": h2, "round2_data_finite": ok,
                    "samples": int(sum(f.shape[0] for f in Fs))})
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
    from the same initial state.  Per substep the RMS residual after the
    iterations is recorded; ``run`` returns the RMS over the substeps per
    configuration and, per learned iteration count k, the number of substeps
    whose residual exceeded the plain run's at the same k and substep (the
    substeps on which the fallback would take over).  The state is restored
    afterwards."""

    def __init__(self, s, theta):
        import torch

        self.s = s
        self.e = _Eager(s)
        self.theta = torch.as_tensor(theta, dtype=torch.float32).to(s.device).contiguous()

    def _keep(self):
        s = self.s
        keep = [t.clone() for t in (s.X, s.V, s.XP, s.LAM)]

        def restore():
            for dst, src in zip((s.X, s.V, s.XP, s.LAM), keep):
                dst.copy_(src)
        return restore

    def run(self, substeps: int, plain_iters=(1, 2, 4), learned_iters=(1, 2), net_every: int = 1):
        """``net_every``: the learned start's network pass runs on every
        ``net_every``-th substep of each frame (the frame plan's
        ``warm_start_every``)."""
        e = self.e
        restore = self._keep()
        runs = {}
        for kind, its in [("plain", k) for k in plain_iters] + [("learned", k)
                                                                 for k in learned_iters]:
            restore()
            r = []
            for k in range(substeps):
                if k % self.s.substeps == 0:
                    e.contacts()
                e.integrate()
                net = (k % self.s.substeps) % max(1, net_every) == 0
                e.iterate(its, theta=self.theta if kind == "learned" else None, net=net)
                r.append(e.residual())
            runs[(kind, its)] = np.asarray(r)
        restore()
        out = {f"{kind}_{k}": float(np.sqrt(np.mean(np.square(v)))) for (kind, k), v in runs.items()}
        fallback = {k: int(np.sum(runs[("learned", k)] > runs[("plain", k)]))
                    for k in learned_iters if ("plain", k) in runs}
        return out, fallback
