"""6-DOF rigid-body XPBD solver — the drop-in for ``xpbd3d.Solver6DOF``
(src/xpbd3d/solver_6dof.py:79-559), on hand-written gfx950 HIP kernels.

Rigid boxes with quaternion orientation, particles (point masses modelled as
small boxes without rotational inertia) and compliant joints share one
substep loop (Mueller et al. 2020, Algorithm 2).  Per frame: the broad phase
runs on the device (``lbvh``: Morton-coded LBVH; ``hashgrid``: counting-sort
uniform grid; ``grid``: the host NumPy grid of the reference) and writes the
candidate pairs plus their count into device buffers; one int is read back
to detect capacity overflow (as the reference does); then the substep loop
is replayed from a hipGraph captured once per plan.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _native as N
from .solver import _resolve_device, grid_candidate_pairs


@dataclass
class RigidBody:
    """Handle of a body (solver_6dof.py:43-50)."""

    index: int
    half_extents: tuple
    color: tuple = (0.6, 0.6, 0.7)
    static: bool = False
    is_particle: bool = False
    group: int = 0


def _world_aabb_half(q, he):
    """|R| he per box for (n,4) xyzw quaternions and (n,3) half extents
    (solver_6dof.py:53-67)."""
    x, y, z, w = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    xx, yy, zz = x * x, y * y, z * z
    xy, xz, yz = x * y, x * z, y * z
    wx, wy, wz = w * x, w * y, w * z
    r = np.abs(np.stack([
        np.stack([1 - 2 * (yy + zz), 2 * (xy - wz), 2 * (xz + wy)], 1),
        np.stack([2 * (xy + wz), 1 - 2 * (xx + zz), 2 * (yz - wx)], 1),
        np.stack([2 * (xz - wy), 2 * (yz + wx), 1 - 2 * (xx + yy)], 1)], 1))
    return np.einsum("nij,nj->ni", r, he).astype(np.float32)


def _box_inv_inertia(he, mass):
    """Diagonal inverse inertia of a solid box (solver_6dof.py:70-76)."""
    lx, ly, lz = 2.0 * he[0], 2.0 * he[1], 2.0 * he[2]
    ix = mass / 12.0 * (ly * ly + lz * lz)
    iy = mass / 12.0 * (lx * lx + lz * lz)
    iz = mass / 12.0 * (lx * lx + ly * ly)
    return (1.0 / ix, 1.0 / iy, 1.0 / iz)


class _Vec3View:
    """Device (n, 3) vector view with a ``.numpy()`` read-back, standing in
    for the reference's ``wp.array(dtype=vec3)`` attributes."""

    def __init__(self, t):
        self._t = t

    def numpy(self):
        return self._t[:, :3].cpu().numpy().copy()

    def __len__(self):
        return self._t.shape[0]


class Solver6DOF:
    """Knobs and defaults of solver_6dof.py:80-97."""

    def __init__(self, dt=1.0 / 60.0, substeps=15, iterations=1, gravity=(0.0, -9.81, 0.0),
                 floor_y=0.0, friction=0.6, restitution=0.0, lin_damp=0.0, ang_damp=0.0,
                 gyroscopic=False, jacobi_relax=1.0, device="cuda:0", max_speed=60.0,
                 max_omega=60.0, broadphase="lbvh"):
        self.device = _resolve_device(device)
        self.dt = float(dt)
        self.substeps = int(substeps)
        self.iterations = int(iterations)
        self.gravity = tuple(float(g) for g in gravity)
        self.floor_y = float(floor_y)
        self.friction = float(friction)
        self.restitution = float(restitution)
        self.lin_damp = float(lin_damp)
        self.ang_damp = float(ang_damp)
        self.gyroscopic = bool(gyroscopic)
        self.jacobi_relax = float(jacobi_relax)
        self.max_speed = float(max_speed)
        self.max_omega = float(max_omega)
        if broadphase not in ("lbvh", "hashgrid", "grid"):
            raise ValueError(f"broadphase {broadphase!r}")
        self.broadphase = str(broadphase)
        self.contact_margin = 0.0
        self._man_margin = 0.02          # manifold clip gap (solver_6dof.py:115-119)
        self.use_graph = True

        self._x, self._q, self._v, self._w = [], [], [], []
        self._inv_mass, self._inv_I, self._he, self._group = [], [], [], []
        self.bodies: list[RigidBody] = []
        self._j_a, self._j_b, self._j_anchor_a, self._j_anchor_b = [], [], [], []
        self._j_rest, self._j_compliance = [], []
        self._dirty = True
        self._buf = None
        self._plan = None
        self._plan_sig = None
        self._graph = None
        self._gscale = None
        self._n_pairs = 0
        self.n_joints = 0
        # "auto": whole frames of LBVH scenes up to BLOCK_AUTO bodies in one
        # workgroup (xpbd6_frame_block); True / False force it on / off
        self.block_frames = "auto"
        # LBVH broad phase: bodies per leaf (1..64), 0 = all pairs, None = the
        # default (all pairs up to 4096 bodies, one body per leaf above);
        # xpbd_amd.autotune.tune6 picks it by timing
        self.lbvh_leaf = None
        self.cap = 0
        self._cur_margin = 0.0

    # ---- scene building (solver_6dof.py:166-229) ----------------------------
    def _add_body(self, position, half_extents, mass, quaternion, velocity, omega, color, static,
                  is_particle, group):
        idx = len(self._x)
        he = tuple(float(h) for h in half_extents)
        self._x.append(tuple(float(p) for p in position))
        self._q.append(tuple(float(c) for c in quaternion))
        self._v.append(tuple(float(c) for c in velocity))
        self._w.append(tuple(float(c) for c in omega))
        if static or mass <= 0.0:
            self._inv_mass.append(0.0)
            self._inv_I.append((0.0, 0.0, 0.0))
            static = True
        elif is_particle:
            self._inv_mass.append(1.0 / float(mass))
            self._inv_I.append((0.0, 0.0, 0.0))
        else:
            self._inv_mass.append(1.0 / float(mass))
            self._inv_I.append(_box_inv_inertia(he, float(mass)))
        self._he.append(he)
        self._group.append(int(group))
        self._dirty = True
        rb = RigidBody(index=idx, half_extents=he, color=color, static=static,
                       is_particle=is_particle, group=int(group))
        self.bodies.append(rb)
        return rb

    def add_box(self, position, half_extents, mass, quaternion=(0.0, 0.0, 0.0, 1.0),
                velocity=(0.0, 0.0, 0.0), omega=(0.0, 0.0, 0.0), color=(0.6, 0.6, 0.75),
                static=False):
        return self._add_body(position, half_extents, mass, quaternion, velocity, omega, color,
                              static, is_particle=False, group=0)

    def add_particle(self, position, mass, radius=0.015, velocity=(0.0, 0.0, 0.0),
                     color=(0.85, 0.3, 0.35), static=False, group=1):
        """Point mass as a small box of half-extent ``radius``; bodies sharing
        a group > 0 do not collide with each other."""
        return self._add_body(position, (radius, radius, radius), mass, (0.0, 0.0, 0.0, 1.0),
                              velocity, (0.0, 0.0, 0.0), color, static, is_particle=True,
                              group=group)

    def add_joint(self, a, b, compliance=0.0, rest_length=None, anchor_a=(0.0, 0.0, 0.0),
                  anchor_b=(0.0, 0.0, 0.0)):
        """Compliant distance joint between body-frame anchors (rest None =
        current anchor separation, measured without rotating the anchors,
        as the reference does, solver_6dof.py:218-222)."""
        ia = a.index if hasattr(a, "index") else int(a)
        ib = b.index if hasattr(b, "index") else int(b)
        aa = np.asarray(anchor_a, np.float64)
        ab = np.asarray(anchor_b, np.float64)
        if rest_length is None:
            xa = np.asarray(self._x[ia])
            xb = np.asarray(self._x[ib])
            rest_length = float(np.linalg.norm((xa + aa) - (xb + ab)))
        self._j_a.append(ia)
        self._j_b.append(ib)
        self._j_anchor_a.append(tuple(float(c) for c in anchor_a))
        self._j_anchor_b.append(tuple(float(c) for c in anchor_b))
        self._j_rest.append(float(rest_length))
        self._j_compliance.append(float(compliance))
        self._dirty = True
        return len(self._j_a) - 1

    def set_gravity_scale(self, scale) -> None:
        """Per-body gravity factor g_i = g * scale[i] (an extension used for
        batches of independent scenes)."""
        self._gscale = np.asarray(scale, np.float32).reshape(-1).copy()
        self._dirty = True

    # ---- host arrays in the reference layout --------------------------------
    def host_arrays(self) -> dict:
        n = len(self._x)
        f = np.float32
        return dict(
            x=np.asarray(self._x, f).reshape(-1, 3), q=np.asarray(self._q, f).reshape(-1, 4),
            v=np.asarray(self._v, f).reshape(-1, 3), w=np.asarray(self._w, f).reshape(-1, 3),
            inv_mass=np.asarray(self._inv_mass, f).reshape(n),
            inv_I=np.asarray(self._inv_I, f).reshape(-1, 3),
            he=np.asarray(self._he, f).reshape(-1, 3),
            group=np.asarray(self._group, np.int32).reshape(n),
            j_a=np.asarray(self._j_a, np.int32), j_b=np.asarray(self._j_b, np.int32),
            j_anchor_a=np.asarray(self._j_anchor_a, f).reshape(-1, 3),
            j_anchor_b=np.asarray(self._j_anchor_b, f).reshape(-1, 3),
            j_rest=np.asarray(self._j_rest, f), j_compliance=np.asarray(self._j_compliance, f))

    # ---- flush (solver_6dof.py:261-354) ---------------------------------------
    def _flush(self):
        if not self._dirty:
            return
        import torch

        N.lib()
        h = self.host_arrays()
        n = h["x"].shape[0]
        if self._buf is not None:
            prev = self._buf["X"].shape[0]
            if 0 < prev <= n:
                h["x"][:prev] = self._buf["X"][:, :3].cpu().numpy()
                h["q"][:prev] = self._buf["Q"].cpu().numpy()
                h["v"][:prev] = self._buf["V"][:, :3].cpu().numpy()
                h["w"][:prev] = self._buf["W"][:, :3].cpu().numpy()
        dev = self.device

        def f4(a3, w=None):
            out = np.zeros((n, 4), np.float32)
            out[:, :3] = a3
            if w is not None:
                out[:, 3] = w
            return torch.from_numpy(out).to(dev)

        he4 = np.zeros((n, 4), np.float32)
        he4[:, :3] = h["he"]
        he4[:, 3] = h["group"].view(np.float32)
        z4 = lambda: torch.zeros((max(n, 1), 4), dtype=torch.float32, device=dev)  # noqa: E731
        b = dict(
            X=f4(h["x"], h["inv_mass"]), Q=torch.from_numpy(h["q"].copy()).to(dev),
            V=f4(h["v"]), W=f4(h["w"]), XP=z4(), QP=z4(), IB=f4(h["inv_I"]),
            HE=torch.from_numpy(he4).to(dev), DX=z4(), DR=z4(), DV=z4(), DW=z4(),
            LAMF=torch.zeros(max(8 * n, 8), dtype=torch.float32, device=dev),
            PC=torch.zeros(1, dtype=torch.int32, device=dev),
            PTC=torch.zeros(1, dtype=torch.int32, device=dev))
        self.n_joints = len(self._j_a)
        nj = self.n_joints
        ji = np.zeros((max(nj, 1), 2), np.int32)
        ja = np.zeros((max(nj, 1), 4), np.float32)
        jb = np.zeros((max(nj, 1), 4), np.float32)
        if nj:
            ji[:, 0], ji[:, 1] = h["j_a"], h["j_b"]
            ja[:nj, :3], ja[:nj, 3] = h["j_anchor_a"], h["j_rest"]
            jb[:nj, :3], jb[:nj, 3] = h["j_anchor_b"], h["j_compliance"]
        b.update(JI=torch.from_numpy(ji).to(dev), JA=torch.from_numpy(ja).to(dev),
                 JB=torch.from_numpy(jb).to(dev),
                 LAMJ=torch.zeros(max(nj, 1), dtype=torch.float32, device=dev))
        if self._gscale is not None:
            if len(self._gscale) != n:
                raise ValueError("gravity scale length != number of bodies")
            b["GS"] = torch.from_numpy(self._gscale).to(dev)
        self._buf = b
        self.cap = max(64, n * 4)
        self._alloc_pair_buffers(self.cap)
        # broad-phase workspace and hash-grid sizing (solver_6dof.py:330-338)
        lib = N.lib()
        ws = max(lib.xpbd6_lbvh_workspace_bytes(n), lib.xpbd6_hashgrid_workspace_bytes(n))
        b["WS"] = torch.empty(int(ws), dtype=torch.uint8, device=dev)
        self._hg_rsmax = float(np.linalg.norm(h["he"], axis=1).max()) if n else 0.0
        self._hg_cell = 2.0 * self._hg_rsmax + self._man_margin
        self._drop_plan()
        self._dirty = False

    def _alloc_pair_buffers(self, cap):
        import torch

        dev = self.device
        b = self._buf
        b["PAIRS"] = torch.zeros((cap, 2), dtype=torch.int32, device=dev)
        b["MC"] = torch.zeros(cap, dtype=torch.int32, device=dev)
        b["PTI"] = torch.zeros((8 * cap, 4), dtype=torch.int32, device=dev)
        b["PTN"] = torch.zeros((8 * cap, 4), dtype=torch.float32, device=dev)
        b["PTOI"] = torch.zeros((8 * cap, 4), dtype=torch.float32, device=dev)
        b["PTOR"] = torch.zeros((8 * cap, 4), dtype=torch.float32, device=dev)
        b["LAMP"] = torch.zeros(8 * cap, dtype=torch.float32, device=dev)

    # one workgroup runs the whole frame up to this many bodies ("auto"):
    # above it the per-launch plan's 256 CUs win the manifold and solve work
    # (box piles, ms per frame, block against launch: 100 bodies 0.50 / 0.67,
    # 150: 0.67 / 0.71, 180: 0.78 / 0.71, 256: 1.03 / 0.71)
    BLOCK_AUTO = 160

    @property
    def n_pairs(self) -> int:
        """Candidate pairs of the last frame (read from the device after a
        whole-frame workgroup launch)."""
        if self._n_pairs is None:
            self._n_pairs = int(self._buf["PC"].item())
        return self._n_pairs

    @n_pairs.setter
    def n_pairs(self, value) -> None:
        self._n_pairs = value

    def _use_block(self, n) -> bool:
        if self.block_frames is False or self.broadphase != "lbvh":
            return False
        if n > int(N.lib().xpbd6_frame_block_capacity()):
            return False
        return self.block_frames is True or n <= self.BLOCK_AUTO

    def _step_block(self, n) -> None:
        import torch

        b = self._buf
        need = max(64, n * (n - 1) // 2)
        if self.cap < need:
            self.cap = need
            self._alloc_pair_buffers(self.cap)
        if "BLO" not in b or b["BLO"].shape[0] != n:
            b["BLO"] = torch.empty((n, 4), dtype=torch.float32, device=self.device)
            b["BHI"] = torch.empty((n, 4), dtype=torch.float32, device=self.device)
        self._get_plan(n)
        N.check(N.lib().xpbd6_frame_block(N.current_stream(), N.ctypes.byref(self._desc),
                                          N.ptr(b["BLO"]), N.ptr(b["BHI"]), self.dt, 1),
                "xpbd6_frame_block")
        self._n_pairs = None

    # ---- plan / graph ---------------------------------------------------------
    def _signature(self, n):
        """Graph key, as solver_6dof.py:481-487 (plus buffer identity)."""
        return (n, self.cap, self.n_joints, self.substeps, self.iterations, self.gyroscopic,
                self.dt, self.gravity, self.floor_y, self.friction, self.restitution,
                self.lin_damp, self.ang_damp, self.max_speed, self.max_omega,
                self._man_margin, self.jacobi_relax, id(self._buf), self._buf["PAIRS"].data_ptr())

    def _drop_plan(self):
        if self._plan is not None:
            N.lib().xpbd_frame6_destroy(self._plan)
        self._plan = None
        self._plan_sig = None

    def _get_plan(self, n):
        sig = self._signature(n)
        if self._plan is not None and self._plan_sig == sig and self._graph is not False:
            return self._plan
        self._drop_plan()
        b = self._buf
        P = N.ptr
        d = N.RigidDesc()
        d.x, d.q, d.v, d.w = P(b["X"]), P(b["Q"]), P(b["V"]), P(b["W"])
        d.x_prev, d.q_prev, d.inv_I, d.he = P(b["XP"]), P(b["QP"]), P(b["IB"]), P(b["HE"])
        d.dx, d.drot, d.dv, d.dw = P(b["DX"]), P(b["DR"]), P(b["DV"]), P(b["DW"])
        d.lam_floor, d.n = P(b["LAMF"]), n
        d.pairs, d.pair_count, d.cap = P(b["PAIRS"]), P(b["PC"]), self.cap
        d.m_count, d.pt_idx = P(b["MC"]), P(b["PTI"])
        d.pt_n, d.pt_oi, d.pt_or = P(b["PTN"]), P(b["PTOI"]), P(b["PTOR"])
        d.lam_pt, d.pt_count = P(b["LAMP"]), P(b["PTC"])
        d.j_idx, d.j_a, d.j_b, d.lam_j = P(b["JI"]), P(b["JA"]), P(b["JB"]), P(b["LAMJ"])
        d.nj = self.n_joints
        d.substeps, d.iterations, d.gyroscopic = self.substeps, self.iterations, int(self.gyroscopic)
        d.h = float(np.float32(self.dt / float(self.substeps)))
        d.gx, d.gy, d.gz = self.gravity
        d.floor_y, d.friction, d.restitution = self.floor_y, self.friction, self.restitution
        d.lin_damp, d.ang_damp, d.relax = self.lin_damp, self.ang_damp, self.jacobi_relax
        d.max_speed, d.max_omega, d.man_margin = self.max_speed, self.max_omega, self._man_margin
        if "GS" in b:
            d.gscale = P(b["GS"])
        self._desc = d
        handle = N.c_void_p()
        N.check(N.lib().xpbd_frame6_create(N.ctypes.byref(d), N.ctypes.byref(handle)),
                "xpbd_frame6_create")
        self._plan = handle
        self._plan_sig = sig
        self._graph = True
        return handle

    def __setattr__(self, name, value):
        # the reference drops its captured graph with `s._graph = None`
        # (tests/test_solver_6dof.py:49); honour that by re-capturing.
        if name == "_graph" and value is None and getattr(self, "_plan", None) is not None:
            object.__setattr__(self, "_graph", False)
            return
        object.__setattr__(self, name, value)

    # ---- broad phase (solver_6dof.py:232-258, 428-479) -------------------------
    def _run_broadphase(self, n, cap):
        b = self._buf
        P = N.ptr
        lib = N.lib()
        s = N.current_stream()
        if self.broadphase == "lbvh" and self.lbvh_leaf is not None:
            rc = lib.xpbd6_broadphase_lbvh_leaf(s, P(b["WS"]), P(b["X"]), P(b["Q"]), P(b["V"]),
                                                P(b["W"]), P(b["HE"]), self._man_margin, self.dt,
                                                n, cap, P(b["PAIRS"]), P(b["PC"]),
                                                int(self.lbvh_leaf))
        elif self.broadphase == "lbvh":
            rc = lib.xpbd6_broadphase_lbvh(s, P(b["WS"]), P(b["X"]), P(b["Q"]), P(b["V"]),
                                           P(b["W"]), P(b["HE"]), self._man_margin, self.dt, n,
                                           cap, P(b["PAIRS"]), P(b["PC"]))
        else:
            rc = lib.xpbd6_broadphase_hashgrid(s, P(b["WS"]), P(b["X"]), P(b["Q"]), P(b["V"]),
                                               P(b["W"]), P(b["HE"]), self._man_margin, self.dt,
                                               self._hg_rsmax, self._hg_cell, n, cap,
                                               P(b["PAIRS"]), P(b["PC"]))
        N.check(rc, f"xpbd6_broadphase_{self.broadphase}")
        return int(b["PC"].item())

    def _grid_pairs(self):
        """Host NumPy grid with a global velocity margin (solver_6dof.py:232-258)."""
        b = self._buf
        pos = b["X"][:, :3].cpu().numpy()
        he = np.asarray(self._he, np.float32).reshape(-1, 3)
        inv_m = np.asarray(self._inv_mass, np.float32)
        n = len(pos)
        if n < 2:
            return np.zeros(0, np.int32), np.zeros(0, np.int32)
        aabb = _world_aabb_half(b["Q"].cpu().numpy().reshape(-1, 4), he)
        vmax = float(np.abs(b["V"][:, :3].cpu().numpy()).max())
        wmax = float(np.abs(b["W"][:, :3].cpu().numpy()).max())
        rmax = float(aabb.max())
        maxhe = float(np.linalg.norm(he, axis=1).max())
        margin = min(self._man_margin + 2.0 * self.dt * (vmax + wmax * maxhe), rmax)
        self._cur_margin = float(margin)
        ga, gb = grid_candidate_pairs(pos, 2.0 * rmax + margin)
        if len(ga) == 0:
            return np.zeros(0, np.int32), np.zeros(0, np.int32)
        sep = np.abs(pos[ga] - pos[gb]) - (aabb[ga] + aabb[gb] + margin)
        keep = np.all(sep <= 0.0, axis=1)
        a = ga[keep].astype(np.int32)
        bb = gb[keep].astype(np.int32)
        grp = np.asarray(self._group, np.int32)
        live = (inv_m[a] > 0.0) | (inv_m[bb] > 0.0)
        live &= ~((grp[a] == grp[bb]) & (grp[a] > 0))
        return a[live], bb[live]

    def step(self):
        self._flush()
        n = len(self._x)
        if n == 0:
            return
        import torch

        b = self._buf
        if self._use_block(n):
            self._step_block(n)
            return
        if self.broadphase in ("lbvh", "hashgrid"):
            cnt = self._run_broadphase(n, self.cap)
            if cnt > self.cap:                       # overflow: grow, re-emit, re-capture
                self.cap = int(cnt * 1.5)
                self._alloc_pair_buffers(self.cap)
                cnt = self._run_broadphase(n, self.cap)
        else:
            a, bb = self._grid_pairs()
            cnt = int(len(a))
            if cnt > self.cap:
                self.cap = int(cnt * 1.5)
                self._alloc_pair_buffers(self.cap)
            if cnt:
                pr = np.stack([a, bb], axis=1).astype(np.int32)
                b["PAIRS"][:cnt] = torch.from_numpy(pr).to(self.device)
            b["PC"].fill_(cnt)
        self.n_pairs = cnt
        plan = self._get_plan(n)
        N.check(N.lib().xpbd_frame6_run(plan, N.current_stream(), 1, int(self.use_graph)),
                "xpbd_frame6_run")

    # ---- perturbation / read-back (solver_6dof.py:532-559) ---------------------
    def set_velocity(self, idx, vel):
        self._flush()
        self._buf["V"][idx, :3] = N_t(vel, self.device)

    def add_impulse(self, idx, dvel):
        self._flush()
        self._buf["V"][idx, :3] += N_t(dvel, self.device)

    @property
    def omega(self):
        self._flush()
        return _Vec3View(self._buf["W"])

    @omega.setter
    def omega(self, value):
        """Accepts an (n,3) array-like (or anything with ``.numpy()``) and
        writes it into the device buffer in place."""
        self._flush()
        arr = value.numpy() if hasattr(value, "numpy") else np.asarray(value)
        self._buf["W"][:, :3] = N_t(np.asarray(arr, np.float32).reshape(-1, 3), self.device)

    def positions(self):
        if self._buf is None:
            return np.asarray(self._x, np.float32).reshape(-1, 3)
        return self._buf["X"][:, :3].cpu().numpy().copy()

    def orientations(self):
        if self._buf is None:
            return np.asarray(self._q, np.float32).reshape(-1, 4)
        return self._buf["Q"].cpu().numpy().copy()

    def velocities(self):
        if self._buf is None:
            return np.asarray(self._v, np.float32).reshape(-1, 3)
        return self._buf["V"][:, :3].cpu().numpy().copy()

    def pairs(self) -> np.ndarray:
        """Current candidate pairs (n_pairs, 2), (i, j) ascending."""
        if self._buf is None:
            return np.zeros((0, 2), np.int32)
        k = min(self.n_pairs, self.cap)
        return self._buf["PAIRS"][:k].cpu().numpy().copy()

    def contact_counts(self) -> np.ndarray:
        """Manifold points per candidate pair (the reference's m_count)."""
        if self._buf is None:
            return np.zeros(0, np.int32)
        return self._buf["MC"][:min(self.n_pairs, self.cap)].cpu().numpy().copy()

    @property
    def num_bodies(self):
        return len(self._x)

    @property
    def num_joints(self):
        return len(self._j_a)

    def __del__(self):
        try:
            self._drop_plan()
        except Exception:
            pass


def N_t(arr, device):
    import torch

    return torch.as_tensor(np.asarray(arr, np.float32)).to(device)
