"""Greedy constraint coloring for the colored Gauss-Seidel sweep.

Same functions and results as src/xpbd3d/coloring.py:26-97, computed by the
library's C++ routine (``xpbd_color_constraints`` / ``xpbd_color_ranges``).
"""
from __future__ import annotations

import numpy as np

from . import _native as N


def color_constraints(n_bodies: int, c_body_a, c_body_b) -> np.ndarray:
    """Color rows so that no two rows of one color share a body
    (coloring.py:26-71).  ``c_body_b[j] < 0`` marks a one-body row.
    Returns int32 colors ``0..k-1`` in row order."""
    a = np.ascontiguousarray(c_body_a, dtype=np.int32)
    b = np.ascontiguousarray(c_body_b, dtype=np.int32)
    n_c = int(a.shape[0])
    color = np.full(n_c, -1, dtype=np.int32)
    if n_c == 0:
        return color
    rc = N.lib().xpbd_color_constraints(int(n_bodies), a.ctypes.data, b.ctypes.data, n_c,
                                        color.ctypes.data)
    if rc < 0:
        raise ValueError("color_constraints: body index out of range")
    return color


def color_ranges(color: np.ndarray) -> tuple[np.ndarray, np.ndarray, int]:
    """Rows sorted by color in CSR form (coloring.py:74-89):
    ``(order, starts, num_colors)`` with color c owning
    ``order[starts[c]:starts[c+1]]``."""
    color = np.ascontiguousarray(color, dtype=np.int32)
    if len(color) == 0:
        return np.zeros(0, dtype=np.int32), np.zeros(1, dtype=np.int32), 0
    num_colors = int(color.max()) + 1
    order = np.empty(len(color), dtype=np.int32)
    starts = np.empty(num_colors + 1, dtype=np.int32)
    rc = N.lib().xpbd_color_ranges(color.ctypes.data, len(color), num_colors, order.ctypes.data,
                                   starts.ctypes.data)
    if rc < 0:
        raise ValueError("color_ranges: invalid color array")
    return order, starts, num_colors


def color_summary(color: np.ndarray) -> dict[int, int]:
    """``{color_id: count}`` (coloring.py:92-97)."""
    out: dict[int, int] = {}
    for c in np.asarray(color).tolist():
        out[int(c)] = out.get(int(c), 0) + 1
    return out
