"""xpbd_amd — XPBD (Macklin et al. 2016; Mueller et al. 2020) on AMD MI355X.

A drop-in for the ``xpbd3d`` package of nanli9/XPBD (src/xpbd3d/__init__.py:7-22):
the same public names, with every kernel hand-written in HIP for gfx950 and
reached through the C-ABI library ``libxpbd_hip.so`` (include/xpbd.h).
"""
from .scene import Body, ConstraintHandle, Shape
from .solver import ATTACH, CONTACT, DISTANCE, FLOOR, Solver

__all__ = [
    "Solver",
    "Body",
    "Shape",
    "ConstraintHandle",
    "DISTANCE",
    "ATTACH",
    "FLOOR",
    "CONTACT",
]

try:  # the 6-DOF solver is optional until its module exists
    from .solver_6dof import RigidBody, Solver6DOF  # noqa: F401

    __all__ += ["Solver6DOF", "RigidBody"]
except ImportError:  # pragma: no cover
    pass
