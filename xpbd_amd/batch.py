"""Many independent scenes of uneven size stepped together (the batched
workload; no reference counterpart — the nearest analogue is several
sub-scenes in one Solver6DOF, examples/viewer_6dof.py:140-213).

Scenes are drawn from a seed: hanging chains of 8-256 links, cloths of
8^2-96^2 particles, sphere piles of 16-2048 spheres of radius 0.05 m and box
piles of 4-256 boxes, sizes log-uniform per type; per scene a gravity scale
U(0.5, 1.5), a compliance log-U(1e-8, 1e-3) and a friction U(0.2, 0.8).

All particle scenes (chains, cloths, sphere piles) share one 3-DOF system
and all box piles one 6-DOF system; ``offsets`` records each scene's index
range (the per-scene offset table).  Every launch of a frame therefore covers
every scene at once (one fused tile-Jacobi launch per iteration over all
rows; sphere CONTACT rows generated on the device each frame), and the
3-DOF and 6-DOF frames run on two streams concurrently.  Scenes sit on a
3 m grid so colliding scenes never touch each other; chains and cloths carry
no collision rows and may overlap freely.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .scenes import cloth_edges

# share of each scene type; with log-uniform sizes the cloths carry most of
# the bodies (a 50/50 mix would average ~600 bodies per scene)
DEFAULT_MIX = {"cloth": 0.90, "pile": 0.04, "chain": 0.03, "boxes": 0.03}
SIZE_RANGE = {"chain": (8, 256), "cloth": (8, 96), "pile": (16, 2048), "boxes": (4, 256)}


@dataclass
class SceneSpec:
    kind: str
    size: int            # links, cloth side, spheres, boxes
    gscale: float
    compliance: float
    mu: float
    origin: tuple
    first: int = 0       # first particle / body index in its system
    count: int = 0
    rows: int = 0


@dataclass
class BatchSpec:
    scenes: list = field(default_factory=list)
    mix: dict = field(default_factory=dict)
    seed: int = 0

    def totals(self):
        out = {"particles": 0, "bodies": 0, "scenes": len(self.scenes)}
        for sc in self.scenes:
            if sc.kind == "boxes":
                out["bodies"] += sc.count
            else:
                out["particles"] += sc.count
        out["total"] = out["particles"] + out["bodies"]
        return out


def draw(n_scenes: int = 1024, seed: int = 0, mix: dict | None = None) -> BatchSpec:
    mix = dict(mix or DEFAULT_MIX)
    kinds = sorted(mix)
    p = np.array([mix[k] for k in kinds], float)
    p /= p.sum()
    rng = np.random.default_rng(seed)
    spec = BatchSpec(mix=dict(zip(kinds, p.tolist())), seed=seed)
    side = int(np.ceil(np.sqrt(n_scenes)))
    for i in range(n_scenes):
        kind = kinds[int(rng.choice(len(kinds), p=p))]
        lo, hi = SIZE_RANGE[kind]
        size = int(round(np.exp(rng.uniform(np.log(lo), np.log(hi)))))
        spec.scenes.append(SceneSpec(
            kind=kind, size=size, gscale=float(rng.uniform(0.5, 1.5)),
            compliance=float(np.exp(rng.uniform(np.log(1e-8), np.log(1e-3)))),
            mu=float(rng.uniform(0.2, 0.8)),
            origin=((i % side) * 3.0, 0.0, (i // side) * 3.0)))
    return spec


class BatchedWorld:
    """Steps a drawn ``BatchSpec``: one 3-DOF Solver (Jacobi, fused tiles)
    holding every particle scene and one Solver6DOF holding every box pile."""

    def __init__(self, spec: BatchSpec, device="cuda:0", substeps=15, iterations=2, dt=1 / 60,
                 seed: int = 0):
        from .solver import Solver
        from .solver_6dof import Solver6DOF

        self.spec = spec
        rng = np.random.default_rng(seed + 1)
        self.particles = Solver(dt=dt, substeps=substeps, iterations=iterations,
                                solve_mode="jacobi", device=device)
        pos, mass, rad, fric, gs = [], [], [], [], []
        ea, eb, rest, comp = [], [], [], []
        fl_a, fl_y, fl_mu = [], [], []
        n_p = 0
        boxes = [sc for sc in spec.scenes if sc.kind == "boxes"]
        for sc in spec.scenes:
            if sc.kind == "boxes":
                continue
            ox, _, oz = sc.origin
            if sc.kind == "chain":
                k = sc.size + 1
                x = np.zeros((k, 3))
                x[:, 0] = ox + 0.05 * np.arange(k)
                x[:, 1] = 2.0 + rng.normal(0, 1e-4, k)
                x[:, 2] = oz
                m = np.ones(k)
                m[0] = 0.0
                a = np.arange(k - 1) + n_p
                ea.append(a), eb.append(a + 1)
                rest.append(np.full(k - 1, 0.05)), comp.append(np.full(k - 1, sc.compliance))
                r = np.zeros(k)
            elif sc.kind == "cloth":
                s_ = sc.size
                k = s_ * s_
                iy, ix = np.divmod(np.arange(k), s_)
                x = np.stack([ox + ix * 0.01, np.full(k, 1.0), oz + iy * 0.01], 1)
                x += rng.normal(0, 1e-4, x.shape)
                m = np.full(k, 1e-3)
                m[0] = m[s_ - 1] = 0.0
                a_l, b_l = cloth_edges(s_)
                ea.append(a_l + n_p), eb.append(b_l + n_p)
                rest.append(np.linalg.norm(x[a_l] - x[b_l], axis=1))
                comp.append(np.full(len(a_l), sc.compliance))
                fl_a.append(np.arange(k) + n_p), fl_y.append(np.zeros(k)), fl_mu.append(np.full(k, sc.mu))
                r = np.zeros(k)
            else:  # sphere pile
                k = sc.size
                side = int(np.ceil(k ** (1 / 3)))
                g = np.arange(k)
                ix, rem = g % side, g // side
                iz, iy = rem % side, rem // side
                x = np.stack([ox + ix * 0.105, 0.05 + iy * 0.105, oz + iz * 0.105], 1)
                x[:, [0, 2]] += rng.uniform(-2e-3, 2e-3, (k, 2))
                m = np.ones(k)
                fl_a.append(np.arange(k) + n_p), fl_y.append(np.full(k, 0.05))
                fl_mu.append(np.full(k, sc.mu))
                r = np.full(k, 0.05)
            sc.first, sc.count = n_p, k
            pos.append(x), mass.append(m), rad.append(r)
            fric.append(np.full(k, sc.mu)), gs.append(np.full(k, sc.gscale))
            n_p += k
        P = self.particles
        if n_p:
            P.add_particles(np.concatenate(pos), np.concatenate(mass), radius=np.concatenate(rad),
                            collide=True, friction=np.concatenate(fric))
            if ea:
                P.add_distances(np.concatenate(ea), np.concatenate(eb), rest=np.concatenate(rest),
                                compliance=np.concatenate(comp))
            if fl_a:
                a = np.concatenate(fl_a)
                y = np.concatenate(fl_y)
                mu = np.concatenate(fl_mu)
                for yv in np.unique(y):
                    sel = y == yv
                    P.add_floor_contacts(a[sel], floor_y=float(yv), friction=mu[sel])
            P.set_gravity_scale(np.concatenate(gs))
            if any(sc.kind == "pile" for sc in spec.scenes):
                P.enable_device_contacts(True)
                # integrate folded into the previous substep's last launch
                # (fuse_first True, the default): 3.27 ms per frame for the
                # particles against 3.54 with a separate integrate pass and
                # 3.76 with integrate folded into each substep's first launch
                # (every contact endpoint then re-reads V)
        self.rigid = None
        if boxes:
            R = Solver6DOF(dt=dt, substeps=substeps, iterations=iterations, device=device,
                           friction=0.5)
            gsb = []
            nb = 0
            for sc in boxes:
                ox, _, oz = sc.origin
                k = sc.size
                side = int(np.ceil(k ** (1 / 3)))
                sc.first = nb
                for j in range(k):
                    ix, rem = j % side, j // side
                    iz, iy = rem % side, rem // side
                    R.add_box((ox + ix * 0.36 + rng.uniform(-1e-3, 1e-3), 0.16 + iy * 0.36,
                               oz + iz * 0.36 + rng.uniform(-1e-3, 1e-3)), (0.16, 0.16, 0.16),
                              mass=1.0)
                sc.count = k
                gsb.append(np.full(k, sc.gscale))
                nb += k
            R.set_gravity_scale(np.concatenate(gsb))
            self.rigid = R
        self._stream_b = None

    def rows(self) -> int:
        """Constraint rows solved per iteration: particle rows + this frame's
        CONTACT rows + 6-DOF rows (joints, manifold points, 8 floor corners
        per dynamic body)."""
        n = self.particles.num_constraints + getattr(self.particles, "n_contacts", 0)
        if self.rigid is not None:
            dyn = int((self.rigid.inv_masses() > 0).sum())
            pts = int(self.rigid.contact_counts().sum()) if self.rigid._buf is not None else 0
            n += self.rigid.num_joints + pts + 8 * dyn
        return n

    def save(self, prefix: str) -> None:
        """Write the batch as scene files (``xpbd_amd.scene_io``):
        ``<prefix>_particles.npz`` and, with box piles, ``<prefix>_rigid.npz``."""
        from . import scene_io

        scene_io.save(self.particles, prefix + "_particles.npz")
        if self.rigid is not None:
            scene_io.save(self.rigid, prefix + "_rigid.npz")

    @classmethod
    def load(cls, prefix: str, device="cuda:0") -> "BatchedWorld":
        """A batch from the scene files written by ``save`` (no spec: the
        per-scene ranges are not stored)."""
        import os

        from . import scene_io

        w = cls.__new__(cls)
        w.spec = None
        w.particles = scene_io.load(prefix + "_particles.npz", device)
        rp = prefix + "_rigid.npz"
        w.rigid = scene_io.load(rp, device) if os.path.exists(rp) else None
        w._stream_b = None
        return w

    def step(self) -> None:
        import torch

        if self.rigid is None:
            self.particles.step()
            return
        if self._stream_b is None:
            self._stream_b = torch.cuda.Stream()
        main = torch.cuda.current_stream()
        # the rigid solver waits for what was queued on the main stream before
        # this step (state writes, the previous frame), not for this step's
        # particle frame: both solvers read a count back once per frame (the
        # contact count, the pair count), each waiting on its own stream only,
        # so the rigid frame runs under the particle frame
        start = torch.cuda.Event()
        start.record(main)
        self.particles.step()
        self._stream_b.wait_event(start)
        with torch.cuda.stream(self._stream_b):
            self.rigid.step()
        main.wait_stream(self._stream_b)
