"""ORACLE — test infrastructure only (the checker, never the product).

ctypes front end of ``liboracle.so``, the serial CPU restatement of the
reference kernels (oracle3.c: src/xpbd3d/kernels.py + coloring.py).  Only
tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg import
this module.

Arrays use the reference's own layout: AoS float32 vec3 (n, 3), float32
scalars, int32 indices.
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
_lib = None

_f = ctypes.c_float
_i = ctypes.c_int
_p = ctypes.c_void_p


def build() -> Path:
    """Compile the oracle with its Makefile (always rebuilds)."""
    subprocess.run(["make", "-B", "-C", str(HERE)], check=True, capture_output=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} missing: run `make -C oracle`")
        lb = ctypes.CDLL(str(LIB))
        lb.orc_integrate.argtypes = [_p, _p, _p, _f, _f, _f, _f, _p, _i]
        lb.orc_solve_color.argtypes = [_p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p,
                                       _i, _i, _f]
        lb.orc_solve_jacobi.argtypes = [_p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _f, _p,
                                        _p, _i]
        lb.orc_apply_jacobi.argtypes = [_p, _p, _p, _p, _f, _i]
        lb.orc_finalize_velocity.argtypes = [_p, _p, _p, _f, _f, _p, _i]
        lb.orc_cap_velocity.argtypes = [_p, _f, _i]
        lb.orc_step3.argtypes = ([_p, _p, _p, _p, _i] + [_p] * 9 + [_i, _p, _p, _i]
                                 + [_f, _f, _f, _f, _i, _i, _i, _f, _f, _f, _i, _p, _p])
        lb.orc_color_constraints.argtypes = [_i, _p, _p, _i, _p]
        lb.orc_color_constraints.restype = _i
        _lib = lb
    return _lib


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def color_constraints(n_bodies: int, a, b) -> np.ndarray:
    """coloring.py:26-71 restated (greedy Welsh-Powell)."""
    a = _c(a, np.int32)
    b = _c(b, np.int32)
    color = np.full(len(a), -1, np.int32)
    if len(a):
        lib().orc_color_constraints(int(n_bodies), a.ctypes.data, b.ctypes.data, len(a),
                                    color.ctypes.data)
    return color


def color_ranges(color):
    """coloring.py:74-89 restated: stable argsort + prefix sums."""
    color = np.asarray(color, np.int32)
    if len(color) == 0:
        return np.zeros(0, np.int32), np.zeros(1, np.int32), 0
    k = int(color.max()) + 1
    order = np.argsort(color, kind="stable").astype(np.int32)
    starts = np.zeros(k + 1, np.int32)
    starts[1:] = np.cumsum(np.bincount(color, minlength=k))
    return order, starts, k


class Scene3:
    """Mutable 3-DOF state in reference layout, stepped by the oracle.

    ``h`` is the scene dict returned by ``Solver.host_arrays()`` (or built by
    hand with the same keys)."""

    def __init__(self, h: dict, *, dt, substeps, iterations, gravity, mode, damping=0.0,
                 relax=1.0, max_speed=40.0):
        self.x = _c(h["x"], np.float32).copy()
        self.v = _c(h["v"], np.float32).copy()
        self.x_prev = np.zeros_like(self.x)
        self.inv_mass = _c(h["inv_mass"], np.float32)
        self.type = _c(h["type"], np.int32)
        self.a = _c(h["a"], np.int32)
        self.b = _c(h["b"], np.int32)
        self.rest = _c(h["rest"], np.float32)
        self.compliance = _c(h["compliance"], np.float32)
        self.anchor = _c(h["anchor"], np.float32).reshape(-1, 3)
        self.mu = _c(h["mu"], np.float32)
        n_c = len(self.type)
        self.lam = np.zeros(n_c, np.float32)
        self.active = np.ones(n_c, np.int32)
        self.dx = np.zeros_like(self.x)
        self.dn = np.zeros(len(self.x), np.int32)
        self.mode = 0 if mode == "gs" else 1
        if self.mode == 0 and n_c:
            color = color_constraints(len(self.x), self.a, self.b)
            self.order, self.starts, self.num_colors = color_ranges(color)
        else:
            self.order = np.zeros(max(n_c, 1), np.int32)
            self.starts = np.array([0, n_c], np.int32)
            self.num_colors = 1
        self.h = float(np.float32(float(dt) / float(substeps)))
        self.substeps = int(substeps)
        self.iterations = int(iterations)
        self.g = tuple(float(v) for v in gravity)
        self.damping = float(damping)
        self.relax = float(relax)
        self.max_speed = float(max_speed)

    def step(self, frames: int = 1) -> None:
        lib().orc_step3(
            self.x.ctypes.data, self.v.ctypes.data, self.x_prev.ctypes.data,
            self.inv_mass.ctypes.data, len(self.x),
            self.type.ctypes.data, self.a.ctypes.data, self.b.ctypes.data,
            self.rest.ctypes.data, self.compliance.ctypes.data, self.anchor.ctypes.data,
            self.mu.ctypes.data, self.lam.ctypes.data, self.active.ctypes.data,
            len(self.type), self.order.ctypes.data, self.starts.ctypes.data, self.num_colors,
            self.g[0], self.g[1], self.g[2], self.h, self.substeps, self.iterations, self.mode,
            self.damping, self.relax, self.max_speed, int(frames),
            self.dx.ctypes.data, self.dn.ctypes.data)


# ---- single-kernel wrappers (arrays are modified in place) -----------------
def integrate(x, v, inv_mass, g, dt, x_prev):
    lib().orc_integrate(x.ctypes.data, v.ctypes.data, inv_mass.ctypes.data, g[0], g[1], g[2],
                        float(dt), x_prev.ctypes.data, len(x))


def _rows(r):
    return [r["type"].ctypes.data, r["a"].ctypes.data, r["b"].ctypes.data,
            r["rest"].ctypes.data, r["compliance"].ctypes.data, r["anchor"].ctypes.data,
            r["mu"].ctypes.data, r["lam"].ctypes.data, r["active"].ctypes.data]


def rows(type, a, b, rest=None, compliance=None, anchor=None, mu=None):
    """Reference-layout row dict with lambda = 0 and every row active."""
    m = len(type)
    z = np.zeros(m, np.float32)
    return dict(
        type=_c(type, np.int32), a=_c(a, np.int32), b=_c(b, np.int32),
        rest=_c(z if rest is None else rest, np.float32),
        compliance=_c(z if compliance is None else compliance, np.float32),
        anchor=_c(np.zeros((m, 3)) if anchor is None else anchor, np.float32).reshape(-1, 3),
        mu=_c(z if mu is None else mu, np.float32),
        lam=np.zeros(m, np.float32), active=np.ones(m, np.int32))


def solve_color(x, x_prev, inv_mass, r, order, start, count, dt):
    order = _c(order, np.int32)
    lib().orc_solve_color(x.ctypes.data, x_prev.ctypes.data, inv_mass.ctypes.data, *_rows(r),
                          order.ctypes.data, int(start), int(count), float(dt))


def solve_jacobi(x, x_prev, inv_mass, r, dt, dx, dn):
    lib().orc_solve_jacobi(x.ctypes.data, x_prev.ctypes.data, inv_mass.ctypes.data, *_rows(r),
                           float(dt), dx.ctypes.data, dn.ctypes.data, len(r["type"]))


def apply_jacobi(x, dx, dn, inv_mass, relax):
    lib().orc_apply_jacobi(x.ctypes.data, dx.ctypes.data, dn.ctypes.data, inv_mass.ctypes.data,
                           float(relax), len(x))


def finalize_velocity(x, x_prev, inv_mass, dt, damping, v):
    lib().orc_finalize_velocity(x.ctypes.data, x_prev.ctypes.data, inv_mass.ctypes.data,
                                float(dt), float(damping), v.ctypes.data, len(x))


def cap_velocity(v, max_speed):
    lib().orc_cap_velocity(v.ctypes.data, float(max_speed), len(v))


# ---------------------------------------------------------------------------
# 6-DOF (oracle6.c: kernels_6dof.py, kernels_bvh.py, kernels_hashgrid.py)
# ---------------------------------------------------------------------------
_SIG6 = {
    "orc6_integrate": [_p] * 6 + [_f, _f, _f, _f, _i, _p, _p, _i],
    "orc6_solve_floor": [_p] * 7 + [_f, _f, _p, _f, _p, _p, _p, _i],
    "orc6_box_manifold": [_p] * 5 + [_i, _f] + [_p] * 7 + [_i],
    "orc6_solve_box": [_p] * 14 + [_i],
    "orc6_solve_joints": [_p] * 11 + [_f, _p, _p, _p, _i],
    "orc6_apply_jacobi": [_p, _p, _p, _f, _p, _p, _p, _i],
    "orc6_finalize_velocity": [_p] * 5 + [_f, _f, _f, _p, _p, _i],
    "orc6_velocity_floor": [_p] * 5 + [_f, _f, _f, _p, _f] + [_p] * 5 + [_i],
    "orc6_velocity_box": [_p] * 10 + [_f, _f, _p, _f] + [_p] * 5 + [_i],
    "orc6_apply_velocity": [_p] * 6 + [_i],
    "orc6_cap_velocity": [_p, _p, _f, _f, _i],
    "orc6_body_aabb": [_p] * 5 + [_f, _f, _p, _p, _i],
    "orc6_pairs_aabb": [_p] * 4 + [_i, _i, _p, _p],
    "orc6_pairs_hashgrid": [_p] * 7 + [_f, _f, _f, _f, _i, _i, _p, _p],
}


def lib6():
    lb = lib()
    if not getattr(lb, "_six", False):
        for name, args in _SIG6.items():
            fn = getattr(lb, name)
            fn.argtypes = args
            fn.restype = _i if name.startswith("orc6_pairs") else None
        lb._six = True
    return lb


def _P(a):
    return a.ctypes.data


class Scene6:
    """6-DOF state in the reference layout (x (n,3), q (n,4) xyzw, v, w,
    inv_mass, inv_I (n,3), he (n,3), group; joints), stepped like
    Solver6DOF.step (solver_6dof.py:489-529): broad phase, then the substep
    sequence of _record (357-426).  Pairs are emitted in (i, j) order."""

    def __init__(self, h: dict, *, dt, substeps, iterations, gravity, floor_y=0.0,
                 friction=0.6, restitution=0.0, lin_damp=0.0, ang_damp=0.0, gyroscopic=False,
                 relax=1.0, max_speed=60.0, max_omega=60.0, broadphase="lbvh",
                 man_margin=0.02):
        f32 = np.float32
        self.x = _c(h["x"], f32).reshape(-1, 3).copy()
        self.q = _c(h["q"], f32).reshape(-1, 4).copy()
        self.v = _c(h["v"], f32).reshape(-1, 3).copy()
        self.w = _c(h["w"], f32).reshape(-1, 3).copy()
        self.inv_mass = _c(h["inv_mass"], f32)
        self.inv_I = _c(h["inv_I"], f32).reshape(-1, 3)
        self.he = _c(h["he"], f32).reshape(-1, 3)
        self.group = _c(h["group"], np.int32)
        n = len(self.x)
        self.n = n
        self.x_prev = np.zeros_like(self.x)
        self.q_prev = np.zeros_like(self.q)
        self.dx = np.zeros_like(self.x)
        self.drot = np.zeros_like(self.x)
        self.dv = np.zeros_like(self.x)
        self.dw = np.zeros_like(self.x)
        self.dcount = np.zeros(n, np.int32)
        self.dvc = np.zeros(n, np.int32)
        self.lam_floor = np.zeros(8 * n, f32)
        self.j_a = _c(h.get("j_a", []), np.int32)
        self.j_b = _c(h.get("j_b", []), np.int32)
        nj = len(self.j_a)
        self.nj = nj
        self.j_anchor_a = _c(h.get("j_anchor_a", np.zeros((nj, 3))), f32).reshape(-1, 3)
        self.j_anchor_b = _c(h.get("j_anchor_b", np.zeros((nj, 3))), f32).reshape(-1, 3)
        self.j_rest = _c(h.get("j_rest", []), f32)
        self.j_compliance = _c(h.get("j_compliance", []), f32)
        self.lam_joint = np.zeros(max(nj, 1), f32)
        self.dt = float(dt)
        self.h = float(np.float32(float(dt) / float(substeps)))
        self.substeps, self.iterations = int(substeps), int(iterations)
        self.g = tuple(float(v) for v in gravity)
        self.floor_y, self.friction, self.restitution = float(floor_y), float(friction), float(restitution)
        self.lin_damp, self.ang_damp = float(lin_damp), float(ang_damp)
        self.gyro = 1 if gyroscopic else 0
        self.relax = float(relax)
        self.max_speed, self.max_omega = float(max_speed), float(max_omega)
        self.broadphase = broadphase
        self.man_margin = float(man_margin)
        self._alloc(max(64, 4 * n))
        self.lowers = np.zeros((max(n, 1), 3), f32)
        self.uppers = np.zeros((max(n, 1), 3), f32)
        self.n_pairs = 0
        # hash-grid cell = 2 * largest bounding-sphere radius + margin
        # (solver_6dof.py:335-337)
        self.rs_max = float(np.linalg.norm(self.he, axis=1).max()) if n else 0.0
        self.cell = 2.0 * self.rs_max + self.man_margin

    def _alloc(self, cap):
        f32 = np.float32
        self.cap = cap
        self.pair_a = np.zeros(cap, np.int32)
        self.pair_b = np.zeros(cap, np.int32)
        self.lam_pair = np.zeros(cap * 8, f32)
        self.m_count = np.zeros(cap, np.int32)
        self.m_inc = np.zeros(cap, np.int32)
        self.m_ref = np.zeros(cap, np.int32)
        self.m_normal = np.zeros((cap, 3), f32)
        self.m_off_inc = np.zeros((cap * 8, 3), f32)
        self.m_off_ref = np.zeros((cap * 8, 3), f32)
        self.poly = np.zeros((cap * 16, 3), f32)

    def broadphase_pairs(self):
        lb = lib6()
        for _ in range(2):
            if self.broadphase == "hashgrid":
                cnt = lb.orc6_pairs_hashgrid(_P(self.x), _P(self.q), _P(self.v), _P(self.w),
                                             _P(self.he), _P(self.inv_mass), _P(self.group),
                                             self.man_margin, self.dt, self.rs_max, self.cell,
                                             self.n, self.cap, _P(self.pair_a),
                                             _P(self.pair_b))
            else:
                lb.orc6_body_aabb(_P(self.x), _P(self.q), _P(self.v), _P(self.w), _P(self.he),
                                  self.man_margin, self.dt, _P(self.lowers), _P(self.uppers),
                                  self.n)
                cnt = lb.orc6_pairs_aabb(_P(self.lowers), _P(self.uppers), _P(self.inv_mass),
                                         _P(self.group), self.n, self.cap, _P(self.pair_a),
                                         _P(self.pair_b))
            if cnt <= self.cap:
                break
            self._alloc(int(cnt * 1.5))
        self.n_pairs = cnt
        return cnt

    def substep(self):
        lb = lib6()
        n, h, cap = self.n, self.h, self.cap
        lb.orc6_integrate(_P(self.x), _P(self.q), _P(self.v), _P(self.w), _P(self.inv_mass),
                          _P(self.inv_I), *self.g, h, self.gyro, _P(self.x_prev),
                          _P(self.q_prev), n)
        self.lam_floor[:] = 0
        self.lam_pair[:] = 0
        self.lam_joint[:] = 0
        lb.orc6_box_manifold(_P(self.x), _P(self.q), _P(self.he), _P(self.pair_a),
                             _P(self.pair_b), self.n_pairs, self.man_margin, _P(self.poly),
                             _P(self.m_count), _P(self.m_inc), _P(self.m_ref),
                             _P(self.m_normal), _P(self.m_off_inc), _P(self.m_off_ref), cap)
        for _ in range(self.iterations):
            if self.nj:
                lb.orc6_solve_joints(_P(self.x), _P(self.q), _P(self.inv_mass), _P(self.inv_I),
                                     _P(self.j_a), _P(self.j_b), _P(self.j_anchor_a),
                                     _P(self.j_anchor_b), _P(self.j_rest),
                                     _P(self.j_compliance), _P(self.lam_joint), h, _P(self.dx),
                                     _P(self.drot), _P(self.dcount), self.nj)
            lb.orc6_solve_floor(_P(self.x), _P(self.q), _P(self.x_prev), _P(self.q_prev),
                                _P(self.inv_mass), _P(self.inv_I), _P(self.he), self.floor_y,
                                self.friction, _P(self.lam_floor), h, _P(self.dx),
                                _P(self.drot), _P(self.dcount), n)
            lb.orc6_solve_box(_P(self.x), _P(self.q), _P(self.inv_mass), _P(self.inv_I),
                              _P(self.m_count), _P(self.m_inc), _P(self.m_ref),
                              _P(self.m_normal), _P(self.m_off_inc), _P(self.m_off_ref),
                              _P(self.lam_pair), _P(self.dx), _P(self.drot), _P(self.dcount),
                              cap)
            lb.orc6_apply_jacobi(_P(self.x), _P(self.q), _P(self.inv_mass), self.relax,
                                 _P(self.dx), _P(self.drot), _P(self.dcount), n)
        lb.orc6_finalize_velocity(_P(self.x), _P(self.q), _P(self.x_prev), _P(self.q_prev),
                                  _P(self.inv_mass), h, self.lin_damp, self.ang_damp,
                                  _P(self.v), _P(self.w), n)
        lb.orc6_velocity_floor(_P(self.x), _P(self.q), _P(self.inv_mass), _P(self.inv_I),
                               _P(self.he), self.floor_y, self.friction, self.restitution,
                               _P(self.lam_floor), h, _P(self.v), _P(self.w), _P(self.dv),
                               _P(self.dw), _P(self.dvc), n)
        lb.orc6_velocity_box(_P(self.x), _P(self.q), _P(self.inv_mass), _P(self.inv_I),
                             _P(self.m_count), _P(self.m_inc), _P(self.m_ref),
                             _P(self.m_normal), _P(self.m_off_inc), _P(self.m_off_ref),
                             self.friction, self.restitution, _P(self.lam_pair), h, _P(self.v),
                             _P(self.w), _P(self.dv), _P(self.dw), _P(self.dvc), cap)
        lb.orc6_apply_velocity(_P(self.inv_mass), _P(self.dv), _P(self.dw), _P(self.dvc),
                               _P(self.v), _P(self.w), n)
        if self.max_speed > 0.0:
            lb.orc6_cap_velocity(_P(self.v), _P(self.w), self.max_speed, self.max_omega, n)

    def step(self, frames: int = 1):
        for _ in range(int(frames)):
            if self.n == 0:
                return
            self.broadphase_pairs()
            for _ in range(self.substeps):
                self.substep()
