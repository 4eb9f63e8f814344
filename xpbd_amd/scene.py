"""Scene-building handles of the public API (src/xpbd3d/scene.py:18-43).

The simulation state lives in device buffers inside the solvers; these
objects only remember indices and carry render metadata.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Literal


@dataclass
class Shape:
    """Visual primitive attached to a body (scene.py:18-26).  The 3-DOF solver
    treats the body as a point mass; for ``sphere`` the first size entry is
    the broad-phase collision radius."""

    kind: Literal["sphere", "cube", "pillar"] = "sphere"
    size: tuple = (0.1,)
    color: tuple = (0.4, 0.6, 0.9)


@dataclass
class Body:
    """Handle to a particle: ``index`` is its row in the body arrays (scene.py:30-35)."""

    index: int
    shape: Shape = field(default_factory=Shape)


@dataclass
class ConstraintHandle:
    """Handle to one or more constraint rows (scene.py:39-43)."""

    index: int
    rows: int = 1
