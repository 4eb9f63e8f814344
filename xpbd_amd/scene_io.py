"""Scene files: one structure-of-arrays ``.npz`` per solver (no pickled
objects: every entry is a plain numeric or unicode array, loadable with
``numpy.load(allow_pickle=False)``).  No reference counterpart: the
reference builds scenes in Python code only (examples/*.py); this format is
the wire format between the scene generators, test fixtures and the batched
world.

A file holds the solver's knobs (``meta_*`` entries), the current state in
API order (positions, velocities; orientations and angular velocities for
rigid bodies) and the static scene (masses, shapes, rows or joints).  Loading
a saved solver and stepping it gives the same trajectory as stepping the
original: the state is stored as the float32 values the device holds, the
scene columns with the precision the solver keeps them in.

Layouts (``kind`` = "xpbd3" for ``Solver``, "xpbd6" for ``Solver6DOF``):

xpbd3  x, v (n,3) float32; mass, radius, friction float64 (n,); collide bool
       (n,); gscale float32 (n,) or absent; rows: type, a, b int32 (m,), rest,
       compliance, mu float32 (m,), anchor float32 (m,3)
xpbd6  x (n,3), q (n,4) xyzw, v, w (n,3) float32; inv_mass float32 (n,),
       inv_I float32 (n,3) (diagonal, body frame), he float32 (n,3), group
       int32 (n,), static bool (n,), particle bool (n,); joints j_a, j_b
       int32, j_anchor_a, j_anchor_b float32 (k,3), j_rest, j_compliance
       float32 (k,); gscale float32 (n,) or absent
"""
from __future__ import annotations

import numpy as np

FORMAT = 1

_KNOBS3 = ("dt", "substeps", "iterations", "damping", "jacobi_relax", "max_speed")
_KNOBS6 = ("dt", "substeps", "iterations", "floor_y", "friction", "restitution", "lin_damp",
           "ang_damp", "jacobi_relax", "max_speed", "max_omega")


def _meta(d: dict, obj, names) -> None:
    for k in names:
        d["meta_" + k] = np.asarray(getattr(obj, k))


def save(solver, path) -> None:
    """Write ``solver`` (a ``Solver`` or ``Solver6DOF``) to ``path`` (.npz)."""
    from .solver import Solver
    from .solver_6dof import Solver6DOF

    d = {"format": np.asarray(FORMAT)}
    if isinstance(solver, Solver):
        d["kind"] = np.asarray("xpbd3")
        _meta(d, solver, _KNOBS3)
        d["meta_gravity"] = np.asarray(solver.gravity, np.float64)
        d["meta_solve_mode"] = np.asarray(solver.solve_mode)
        d["meta_self_collision"] = np.asarray(bool(solver._self_collision))
        d["meta_contact_margin"] = np.asarray(float(solver._contact_margin))
        d["meta_device_contacts"] = np.asarray(bool(solver._device_contacts))
        d["meta_default_friction"] = np.asarray(float(solver._default_friction))
        d["x"] = solver.positions().astype(np.float32)
        d["v"] = solver.velocities().astype(np.float32)
        d["mass"] = solver._mass.array().astype(np.float64)
        d["radius"] = solver._radius.array().astype(np.float64)
        d["friction"] = solver._friction.array().astype(np.float64)
        d["collide"] = solver._collide.array().astype(bool)
        keep = solver._rt.array() != 3 if solver._self_collision else slice(None)
        d["type"] = solver._rt.array().astype(np.int32)[keep]
        d["a"] = solver._ra.array().astype(np.int32)[keep]
        d["b"] = solver._rb.array().astype(np.int32)[keep]
        d["rest"] = solver._rrest.array().astype(np.float32)[keep]
        d["compliance"] = solver._rcomp.array().astype(np.float32)[keep]
        d["mu"] = solver._rmu.array().astype(np.float32)[keep]
        d["anchor"] = solver._ranch.array().astype(np.float32).reshape(-1, 3)[keep]
        if solver._gscale is not None:
            d["gscale"] = np.asarray(solver._gscale, np.float32)
    elif isinstance(solver, Solver6DOF):
        d["kind"] = np.asarray("xpbd6")
        _meta(d, solver, _KNOBS6)
        d["meta_gravity"] = np.asarray(solver.gravity, np.float64)
        d["meta_gyroscopic"] = np.asarray(bool(solver.gyroscopic))
        d["meta_broadphase"] = np.asarray(solver.broadphase)
        h = solver.host_arrays()
        if solver._buf is not None and not solver._dirty:
            h["x"] = solver.positions().astype(np.float32)
            h["q"] = solver.orientations().astype(np.float32)
            h["v"] = solver.velocities().astype(np.float32)
            h["w"] = np.asarray(solver.omega.numpy(), np.float32).reshape(-1, 3)
        for k in ("x", "q", "v", "w", "inv_mass", "inv_I", "he", "group", "j_a", "j_b",
                  "j_anchor_a", "j_anchor_b", "j_rest", "j_compliance"):
            d[k] = h[k]
        d["static"] = np.asarray([b.static for b in solver.bodies], bool)
        d["particle"] = np.asarray([b.is_particle for b in solver.bodies], bool)
        d["color"] = np.asarray([b.color for b in solver.bodies], np.float64).reshape(-1, 3)
        if solver._gscale is not None:
            d["gscale"] = np.asarray(solver._gscale, np.float32)
    else:
        raise TypeError(f"cannot save {type(solver).__name__}")
    np.savez(path, **d)


def load(path, device="cuda:0"):
    """Read a scene file written by ``save``; returns a new solver on
    ``device``."""
    with np.load(path, allow_pickle=False) as f:
        d = {k: f[k] for k in f.files}
    if int(d["format"]) != FORMAT:
        raise ValueError(f"scene format {int(d['format'])}, expected {FORMAT}")
    kind = str(d["kind"])
    if kind == "xpbd3":
        return _load3(d, device)
    if kind == "xpbd6":
        return _load6(d, device)
    raise ValueError(f"unknown scene kind {kind!r}")


def _load3(d, device):
    from .solver import Solver

    s = Solver(dt=float(d["meta_dt"]), substeps=int(d["meta_substeps"]),
               iterations=int(d["meta_iterations"]),
               gravity=tuple(float(g) for g in d["meta_gravity"]),
               damping=float(d["meta_damping"]), solve_mode=str(d["meta_solve_mode"]),
               jacobi_relax=float(d["meta_jacobi_relax"]), device=device,
               max_speed=float(d["meta_max_speed"]))
    s._default_friction = float(d["meta_default_friction"])
    s._x.extend(d["x"].astype(np.float64))
    s._v.extend(d["v"].astype(np.float64))
    s._mass.extend(d["mass"])
    s._radius.extend(d["radius"])
    s._collide.extend(d["collide"])
    s._friction.extend(d["friction"])
    s._rt.extend(d["type"])
    s._ra.extend(d["a"])
    s._rb.extend(d["b"])
    s._rrest.extend(d["rest"])
    s._rcomp.extend(d["compliance"])
    s._rmu.extend(d["mu"])
    s._ranch.extend(d["anchor"])
    if "gscale" in d:
        s.set_gravity_scale(d["gscale"])
    if bool(d["meta_self_collision"]):
        s.enable_self_collision(True, margin=float(d["meta_contact_margin"]))
    if bool(d["meta_device_contacts"]):
        s.enable_device_contacts(True, margin=float(d["meta_contact_margin"]))
    s._dirty = True
    return s


def _load6(d, device):
    from .solver_6dof import _POINT, _STATIC, Solver6DOF

    s = Solver6DOF(dt=float(d["meta_dt"]), substeps=int(d["meta_substeps"]),
                   iterations=int(d["meta_iterations"]),
                   gravity=tuple(float(g) for g in d["meta_gravity"]),
                   floor_y=float(d["meta_floor_y"]), friction=float(d["meta_friction"]),
                   restitution=float(d["meta_restitution"]), lin_damp=float(d["meta_lin_damp"]),
                   ang_damp=float(d["meta_ang_damp"]), gyroscopic=bool(d["meta_gyroscopic"]),
                   jacobi_relax=float(d["meta_jacobi_relax"]), device=device,
                   max_speed=float(d["meta_max_speed"]), max_omega=float(d["meta_max_omega"]),
                   broadphase=str(d["meta_broadphase"]))
    # bodies as stored (inverse masses and inertias taken from the file,
    # not recomputed), then the joints with their stored rest lengths
    kind = (np.where(d["static"], _STATIC, 0) | np.where(d["particle"], _POINT, 0))
    color = d["color"] if "color" in d else np.full((d["x"].shape[0], 3), 0.6)
    s._scene.append_bodies(dict(x=d["x"], q=d["q"], v=d["v"], w=d["w"],
                                inv_mass=d["inv_mass"], inv_I=d["inv_I"], he=d["he"],
                                color=color), d["group"], kind.astype(np.int8))
    s.add_joints(d["j_a"], d["j_b"], compliance=d["j_compliance"], rest_length=d["j_rest"],
                 anchor_a=d["j_anchor_a"], anchor_b=d["j_anchor_b"])
    if "gscale" in d:
        s.set_gravity_scale(d["gscale"])
    s._dirty = True
    return s
