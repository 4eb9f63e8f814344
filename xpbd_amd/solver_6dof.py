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
from .solver import _Col, _resolve_device, grid_candidate_pairs


@dataclass
class RigidBody:
    """Handle of one body, as ``xpbd3d.RigidBody`` (solver_6dof.py:43-50):
    its index, half extents, display colour, and whether it is static, a
    point mass, and its collision group."""

    index: int
    half_extents: tuple
    color: tuple = (0.6, 0.6, 0.7)
    static: bool = False
    is_particle: bool = False
    group: int = 0


# bit flags of the per-body ``kind`` column
_STATIC = 1
_POINT = 2


def solid_box_inverse_inertia(half_extents, mass) -> np.ndarray:
    """Body-frame diagonal inverse inertia of solid boxes, (n, 3): with edge
    lengths l = 2 he, I_k = m (|l|^2 - l_k^2) / 12 (the two other edges)."""
    he = np.asarray(half_extents, np.float64).reshape(-1, 3)
    m = np.asarray(mass, np.float64).reshape(-1, 1)
    l2 = (2.0 * he) ** 2
    return 1.0 / (m / 12.0 * (l2.sum(axis=1, keepdims=True) - l2))


def rotation_matrices(q) -> np.ndarray:
    """Rotation matrices (n, 3, 3) of unit quaternions (n, 4) stored x, y, z, w."""
    q = np.asarray(q).reshape(-1, 4)
    v, w = q[:, :3], q[:, 3]
    # R = (w^2 - |v|^2) I + 2 v v^T + 2 w [v]x
    outer = 2.0 * v[:, :, None] * v[:, None, :]
    diag = (w * w - (v * v).sum(axis=1))[:, None, None] * np.eye(3, dtype=q.dtype)
    x, y, z = v[:, 0], v[:, 1], v[:, 2]
    zero = np.zeros_like(x)
    skew = np.stack([np.stack([zero, -z, y], -1), np.stack([z, zero, -x], -1),
                     np.stack([-y, x, zero], -1)], 1)
    return diag + outer + 2.0 * w[:, None, None] * skew


def world_aabb_half(q, he) -> np.ndarray:
    """World-AABB half extents |R| he of oriented boxes, float32 (n, 3)."""
    R = rotation_matrices(np.asarray(q, np.float32))
    return (np.abs(R) @ np.asarray(he, np.float32).reshape(-1, 3, 1))[:, :, 0].astype(np.float32)


def host_grid_pairs(x, q, he, v, w, inv_mass, group, dt, man_margin):
    """Candidate pairs of the host ``grid`` broad phase (solver_6dof.py:232-258):
    one global margin m = min(man_margin + 2 dt (max|v| + max|w| max|he|),
    r_max) over the float32 world AABBs, a uniform grid of cell 2 r_max + m
    enumerated as the 3-DOF contact grid is, then the per-axis overlap test
    |dx| <= a_i + a_j + m and the filters (not both static, not the same
    collision group > 0).  Returns (a, b, m)."""
    x = np.asarray(x, np.float32).reshape(-1, 3)
    n = x.shape[0]
    none = np.zeros(0, np.int32)
    if n < 2:
        return none, none, 0.0
    he = np.asarray(he, np.float32).reshape(-1, 3)
    half = world_aabb_half(q, he)
    r_max = float(half.max())
    speed = float(np.abs(np.asarray(v, np.float32)).max())
    spin = float(np.abs(np.asarray(w, np.float32)).max())
    reach = float(np.linalg.norm(he, axis=1).max())
    margin = min(man_margin + 2.0 * dt * (speed + spin * reach), r_max)
    ca, cb = grid_candidate_pairs(x, 2.0 * r_max + margin)
    if len(ca) == 0:
        return none, none, float(margin)
    gap = np.abs(x[ca] - x[cb]) - (half[ca] + half[cb] + margin)
    inv_m = np.asarray(inv_mass, np.float32)
    grp = np.asarray(group, np.int32)
    ok = (gap <= 0.0).all(axis=1)
    ok &= (inv_m[ca] > 0.0) | (inv_m[cb] > 0.0)
    ok &= (grp[ca] != grp[cb]) | (grp[ca] <= 0)
    return ca[ok].astype(np.int32), cb[ok].astype(np.int32), float(margin)


class _Vec3View:
    """Device (n, 3) vector view with a ``.numpy()`` read-back, standing in
    for the reference's ``wp.array(dtype=vec3)`` attributes."""

    def __init__(self, t):
        self._t = t

    def numpy(self):
        return self._t[:, :3].cpu().numpy().copy()

    def __len__(self):
        return self._t.shape[0]


class _RigidScene:
    """The host scene as columns (float64, rounded to float32 once at
    upload): bodies (pose, velocities, inverse mass and inertia, half
    extents, collision group, static/point-mass flags, colour) and joints.
    Single adds and bulk adds append to the same columns."""

    _BODY = (("x", 3), ("q", 4), ("v", 3), ("w", 3), ("inv_mass", 0), ("inv_I", 3),
             ("he", 3), ("color", 3))
    _JOINT = (("anchor_a", 3), ("anchor_b", 3), ("rest", 0), ("compliance", 0))

    def __init__(self):
        self.body = {k: _Col(np.float64, w) for k, w in self._BODY}
        self.group = _Col(np.int32)
        self.kind = _Col(np.int8)
        self.joint = {k: _Col(np.float64, w) for k, w in self._JOINT}
        self.ja = _Col(np.int32)
        self.jb = _Col(np.int32)

    def n_bodies(self) -> int:
        return len(self.group)

    def n_joints(self) -> int:
        return len(self.ja)

    def append_bodies(self, cols: dict, group, kind) -> np.ndarray:
        base = self.n_bodies()
        for k, _ in self._BODY:
            self.body[k].extend(cols[k])
        self.group.extend(group)
        self.kind.extend(kind)
        return np.arange(base, self.n_bodies(), dtype=np.int64)

    def append_joints(self, a, b, cols: dict) -> np.ndarray:
        base = self.n_joints()
        self.ja.extend(a)
        self.jb.extend(b)
        for k, _ in self._JOINT:
            self.joint[k].extend(cols[k])
        return np.arange(base, self.n_joints(), dtype=np.int64)

    def arrays(self) -> dict:
        """float32 / int32 arrays in the reference's layout."""
        f = np.float32
        b = {k: c.array().astype(f) for k, c in self.body.items()}
        j = {k: c.array().astype(f) for k, c in self.joint.items()}
        return dict(x=b["x"], q=b["q"], v=b["v"], w=b["w"], inv_mass=b["inv_mass"],
                    inv_I=b["inv_I"], he=b["he"], group=self.group.array().astype(np.int32),
                    j_a=self.ja.array().astype(np.int32), j_b=self.jb.array().astype(np.int32),
                    j_anchor_a=j["anchor_a"], j_anchor_b=j["anchor_b"], j_rest=j["rest"],
                    j_compliance=j["compliance"])


def _rows(a, n, width):
    """Broadcast a per-body value (scalar, row or (n, width) array) to (n, width)."""
    return np.broadcast_to(np.asarray(a, np.float64), (n, width) if width else (n,))


def _index(body) -> int:
    return int(body.index) if hasattr(body, "index") else int(body)


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

        self._scene = _RigidScene()
        self._handles: list = []
        self._dirty = True
        self._dev = None                 # device buffers (read them through _buf)
        self._bp_in_plan = False         # the plan runs the broad phase itself
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
        # launches per substep of the frame plan: 2 (per-body phases folded
        # into the contact and velocity launches, xpbd6_substep_*) or 5 (one
        # launch per phase); the same results up to the order of atomic sums
        self.substep_launches = 2
        self.cap = 0
        self._cur_margin = 0.0

    # ---- scene building (solver_6dof.py:166-229) ----------------------------
    def add_boxes(self, positions, half_extents, masses, quaternions=(0.0, 0.0, 0.0, 1.0),
                  velocities=(0.0, 0.0, 0.0), omegas=(0.0, 0.0, 0.0), color=(0.6, 0.6, 0.75),
                  static=False, particle=False, group=0) -> np.ndarray:
        """Bulk body creation (an extension; ``add_box`` / ``add_particle``
        add one).  Every argument is one value or one per body.  A body is
        static (inverse mass and inertia 0) if ``static`` or its mass is not
        positive; point masses (``particle``) get no rotational inertia;
        other bodies the inertia of a solid box.  Returns the new indices."""
        x = np.asarray(positions, np.float64).reshape(-1, 3)
        n = x.shape[0]
        he = _rows(half_extents, n, 3)
        mass = _rows(masses, n, 0)
        stat = _rows(static, n, 0).astype(bool) | (mass <= 0.0)
        point = _rows(particle, n, 0).astype(bool)
        dyn = ~stat
        inv_mass = np.zeros(n)
        inv_mass[dyn] = 1.0 / mass[dyn]
        inv_I = np.zeros((n, 3))
        spin = dyn & ~point
        if spin.any():
            inv_I[spin] = solid_box_inverse_inertia(he[spin], mass[spin])
        cols = dict(x=x, q=_rows(quaternions, n, 4), v=_rows(velocities, n, 3),
                    w=_rows(omegas, n, 3), inv_mass=inv_mass, inv_I=inv_I, he=he,
                    color=_rows(color, n, 3))
        kind = np.where(stat, _STATIC, 0) | np.where(point, _POINT, 0)
        idx = self._scene.append_bodies(cols, _rows(group, n, 0).astype(np.int32),
                                        kind.astype(np.int8))
        self._dirty = True
        return idx

    def _handle(self, i: int) -> RigidBody:
        sc = self._scene
        kind = int(sc.kind[i])
        return RigidBody(index=i, half_extents=tuple(float(c) for c in sc.body["he"][i]),
                         color=tuple(float(c) for c in sc.body["color"][i]),
                         static=bool(kind & _STATIC), is_particle=bool(kind & _POINT),
                         group=int(sc.group[i]))

    def add_box(self, position, half_extents, mass, quaternion=(0.0, 0.0, 0.0, 1.0),
                velocity=(0.0, 0.0, 0.0), omega=(0.0, 0.0, 0.0), color=(0.6, 0.6, 0.75),
                static=False) -> RigidBody:
        """One solid box (group 0: collides with everything)."""
        i = int(self.add_boxes([position], half_extents, mass, quaternion, velocity, omega,
                               color, static)[0])
        return self._handle(i)

    def add_particle(self, position, mass, radius=0.015, velocity=(0.0, 0.0, 0.0),
                     color=(0.85, 0.3, 0.35), static=False, group=1) -> RigidBody:
        """A point mass: a box of half extent ``radius`` for contacts, no
        rotational inertia; bodies sharing a group > 0 do not collide with
        each other (cloth nodes)."""
        i = int(self.add_boxes([position], (radius, radius, radius), mass,
                               velocities=velocity, color=color, static=static, particle=True,
                               group=group)[0])
        return self._handle(i)

    def add_joints(self, a, b, compliance=0.0, rest_length=None, anchor_a=(0.0, 0.0, 0.0),
                   anchor_b=(0.0, 0.0, 0.0)) -> np.ndarray:
        """Bulk ``add_joint`` over index arrays (an extension)."""
        a = np.asarray(a, np.int64).reshape(-1)
        b = np.asarray(b, np.int64).reshape(-1)
        m = a.shape[0]
        aa, ab = _rows(anchor_a, m, 3), _rows(anchor_b, m, 3)
        if rest_length is None:
            # the anchors' separation at the bodies' current positions,
            # anchors taken unrotated (solver_6dof.py:218-222)
            x = self._scene.body["x"].array()
            rest = np.linalg.norm((x[a] + aa) - (x[b] + ab), axis=1)
        else:
            rest = _rows(rest_length, m, 0)
        idx = self._scene.append_joints(a, b, dict(anchor_a=aa, anchor_b=ab, rest=rest,
                                                   compliance=_rows(compliance, m, 0)))
        self._dirty = True
        return idx

    def add_joint(self, a, b, compliance=0.0, rest_length=None, anchor_a=(0.0, 0.0, 0.0),
                  anchor_b=(0.0, 0.0, 0.0)) -> int:
        """Compliant distance joint between body-frame anchors of bodies (or
        indices) ``a`` and ``b``; ``rest_length=None`` = their current
        separation, ``compliance`` 0 = rigid.  Returns the joint index."""
        return int(self.add_joints([_index(a)], [_index(b)], compliance, rest_length, anchor_a,
                                   anchor_b)[0])

    @property
    def bodies(self) -> list:
        """RigidBody handles of every body (built from the columns)."""
        n = self._scene.n_bodies()
        if len(self._handles) != n:
            self._handles = self._handles[:n] + [self._handle(i)
                                                 for i in range(len(self._handles), n)]
        return self._handles

    def inv_masses(self) -> np.ndarray:
        """Inverse masses (float32, API order); 0 marks static bodies."""
        return self._scene.body["inv_mass"].array().astype(np.float32)

    def set_gravity_scale(self, scale) -> None:
        """Per-body gravity factor g_i = g * scale[i] (an extension used for
        batches of independent scenes)."""
        self._gscale = np.asarray(scale, np.float32).reshape(-1).copy()
        self._dirty = True

    # ---- host arrays in the reference layout --------------------------------
    def host_arrays(self) -> dict:
        return self._scene.arrays()

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
            PTC=torch.zeros(2, dtype=torch.int32, device=dev),
            AUX=torch.zeros((6 * max(n, 1), 4), dtype=torch.float32, device=dev))
        self.n_joints = self._scene.n_joints()
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
        self._dev = b
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
        b = self._dev
        b["PAIRS"] = torch.zeros((cap, 2), dtype=torch.int32, device=dev)
        b["MC"] = torch.zeros(cap, dtype=torch.int32, device=dev)
        b["PTI"] = torch.zeros((8 * cap, 4), dtype=torch.int32, device=dev)
        b["PTN"] = torch.zeros((8 * cap, 4), dtype=torch.float32, device=dev)
        b["PTOI"] = torch.zeros((8 * cap, 4), dtype=torch.float32, device=dev)
        b["PTOR"] = torch.zeros((8 * cap, 4), dtype=torch.float32, device=dev)
        b["LAMP"] = torch.zeros(8 * cap, dtype=torch.float32, device=dev)

    # one workgroup runs the whole frame up to this many bodies ("auto"):
    # above it the two-launch plan's 256 CUs win the manifold and solve work
    # (box piles, ms per frame, block against two-launch plan: 27 bodies
    # 0.308 / 0.318, 100: 0.498 / 0.365, 144: 0.624 / 0.389, 256: 1.03 / 0.41)
    BLOCK_AUTO = 32

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

        b = self._dev
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
                self._man_margin, self.jacobi_relax, id(self._dev), self._dev["PAIRS"].data_ptr(),
                self.substep_launches, self.broadphase, self.lbvh_leaf)

    def _drop_plan(self):
        if self._plan is not None:
            self._settle_frames()
            N.lib().xpbd_frame6_destroy(self._plan)
        self._plan = None
        self._plan_sig = None
        self._bp_in_plan = False

    # ---- broad phase inside the frame plan -----------------------------------
    def _plan_runs_broadphase(self) -> bool:
        """LBVH / hash-grid frames of the two-launch plan run their broad
        phase inside the plan: no host round trip per frame."""
        return (self.broadphase in ("lbvh", "hashgrid") and self.substep_launches == 2
                and self.iterations >= 1)

    def _status(self, wait: bool):
        out = (N.ctypes.c_longlong * 4)()
        N.check(N.lib().xpbd_frame6_status(self._plan, int(wait), out),
                "xpbd_frame6_status")
        return [int(v) for v in out]

    def _settle_frames(self, wait: bool = True) -> None:
        """Reads the status of the frames the plan has run (all of them when
        ``wait``).  A frame whose broad phase found more pairs than the pair
        buffers hold wrote nothing, and neither did any frame after it (the
        state did not change, so neither did their pairs): grow the buffers
        as solver_6dof.py:443-453 does and run those frames again, each
        checked before the next."""
        if self._plan is None or not self._bp_in_plan:
            return
        _, bad, count, redo = self._status(wait)
        if bad < 0:
            return
        self._status(True)                 # every skipped frame has ended
        plan = self._plan
        self._plan = None                  # its frames are accounted for here
        N.lib().xpbd_frame6_destroy(plan)
        self._plan_sig = None
        self._bp_in_plan = False
        n = self.num_bodies
        k = 0
        while k < redo:
            if count > self.cap:
                self.cap = int(count * 1.5)
                self._alloc_pair_buffers(self.cap)
            plan = self._get_plan(n)
            N.check(N.lib().xpbd_frame6_run(plan, N.current_stream(), 1, int(self.use_graph)),
                    "xpbd_frame6_run")
            _, bad, count, _ = self._status(True)
            if bad < 0:
                k += 1
            else:                          # grew again: drop the plan, same frame again
                p = self._plan
                self._plan = None
                N.lib().xpbd_frame6_destroy(p)
                self._plan_sig = None
                self._bp_in_plan = False

    def settle(self) -> None:
        """Waits for the frames run so far and makes sure every one of them
        ran (see _settle_frames); reads of the device state do this first."""
        self._settle_frames(True)

    @property
    def _buf(self):
        """The device buffers, after every frame run so far has ended and
        been checked (None before the first flush)."""
        if self.__dict__.get("_plan") is not None and self.__dict__.get("_bp_in_plan"):
            self._settle_frames(True)
        return self.__dict__.get("_dev")

    def _get_plan(self, n):
        sig = self._signature(n)
        if self._plan is not None and self._plan_sig == sig and self._graph is not False:
            return self._plan
        self._drop_plan()
        b = self._dev
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
        if self.substep_launches not in (2, 5):
            raise ValueError("substep_launches must be 2 or 5")
        if self.substep_launches == 2:
            d.aux = P(b["AUX"])
        self._desc = d
        handle = N.c_void_p()
        N.check(N.lib().xpbd_frame6_create(N.ctypes.byref(d), N.ctypes.byref(handle)),
                "xpbd_frame6_create")
        self._plan = handle
        self._plan_sig = sig
        self._graph = True
        if self._plan_runs_broadphase() and not self._use_block(n):
            kind = 1 if self.broadphase == "lbvh" else 2
            leaf = -1 if self.lbvh_leaf is None else int(self.lbvh_leaf)
            N.check(N.lib().xpbd_frame6_set_broadphase(handle, kind, P(b["WS"]), leaf,
                                                        self._hg_rsmax, self._hg_cell, self.dt),
                    "xpbd_frame6_set_broadphase")
            self._bp_in_plan = True
        return handle

    # knobs the frame plan is built from: a change first settles the frames
    # run with the old values (a skipped frame runs again with them)
    _PLAN_KNOBS = frozenset((
        "dt", "substeps", "iterations", "gravity", "floor_y", "friction", "restitution",
        "lin_damp", "ang_damp", "gyroscopic", "jacobi_relax", "max_speed", "max_omega",
        "broadphase", "lbvh_leaf", "substep_launches", "use_graph", "block_frames",
        "_man_margin", "_gscale"))

    def __setattr__(self, name, value):
        # the reference drops its captured graph with `s._graph = None`
        # (tests/test_solver_6dof.py:49); honour that by re-capturing.
        if name == "_graph" and value is None and getattr(self, "_plan", None) is not None:
            object.__setattr__(self, "_graph", False)
            return
        if name in self._PLAN_KNOBS and self.__dict__.get("_bp_in_plan"):
            self._settle_frames(True)
        object.__setattr__(self, name, value)

    # ---- broad phase (solver_6dof.py:232-258, 428-479) -------------------------
    def _run_broadphase(self, n, cap):
        b = self._dev
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
        """The host ``grid`` broad phase on the device state (read back)."""
        b = self._buf
        h = self.host_arrays()
        a, bb, self._cur_margin = host_grid_pairs(
            b["X"][:, :3].cpu().numpy(), b["Q"].cpu().numpy(), h["he"],
            b["V"][:, :3].cpu().numpy(), b["W"][:, :3].cpu().numpy(), h["inv_mass"],
            h["group"], self.dt, self._man_margin)
        return a, bb

    def step(self):
        self._flush()
        n = self.num_bodies
        if n == 0:
            return
        import torch

        b = self._dev
        if self._use_block(n):
            self._step_block(n)
            return
        if self._plan_runs_broadphase():
            self._settle_frames(False)     # only what the device has reported
            plan = self._get_plan(n)
            N.check(N.lib().xpbd_frame6_run(plan, N.current_stream(), 1, int(self.use_graph)),
                    "xpbd_frame6_run")
            self._n_pairs = None
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
            return self.host_arrays()["x"]
        return self._buf["X"][:, :3].cpu().numpy().copy()

    def orientations(self):
        if self._buf is None:
            return self.host_arrays()["q"]
        return self._buf["Q"].cpu().numpy().copy()

    def velocities(self):
        if self._buf is None:
            return self.host_arrays()["v"]
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
        return self._scene.n_bodies()

    @property
    def num_joints(self):
        return self._scene.n_joints()

    def __del__(self):
        try:
            self._drop_plan()
        except Exception:
            pass


def N_t(arr, device):
    import torch

    return torch.as_tensor(np.asarray(arr, np.float32)).to(device)
