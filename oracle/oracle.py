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
