"""ORACLE — test infrastructure only.  NumPy restatement of the learned
warm-start row computations in float64: the row features (checked against
``xpbd_ws_features``), the warm-start launch (``xpbd_jacobi_tile_warm``,
given the network outputs) and the residual (``xpbd_ws_residual``).  No
reference counterpart: the warm start is new; the row geometry follows the
projection conventions of src/xpbd3d/kernels.py:131-208 (C = |x_a - x_b| -
rest with the normal from b to a for two-body rows, C = y - floor_y with the
normal +y for FLOOR rows) and the Jacobi averaging of kernels.py:326-339.

Layouts: X, XP float (n, 4) (x, y, z, inverse mass; XP's 4th column
unused); rows int (m, 4) packed as in include/xpbd.h; anchors float (m, 4).
Velocities are (X - XP) / h.
"""
from __future__ import annotations

import numpy as np

LEN = 1e-3
OMEGA_MAX = 4.0
EPS = 1e-9
DISTANCE, ATTACH, FLOOR, CONTACT = 0, 1, 2, 3


def slog(x):
    return np.sign(x) * np.log1p(np.abs(x))


def _unpack(rows):
    rows = np.asarray(rows, np.int32)
    t = (rows[:, 0].view(np.uint32) >> 28).astype(np.int64)
    a = (rows[:, 0] & 0x0FFFFFFF).astype(np.int64)
    b = rows[:, 1].astype(np.int64)
    p0 = rows[:, 2].view(np.float32).astype(np.float64)
    p1 = rows[:, 3].view(np.float32).astype(np.float64)
    return t, a, b, p0, p1


def geometry(X, XP, rows, anchors, h):
    """Per row: type, a, b, C, w, alpha~, normal (m, 3), v_n, |v_t|,
    penetration, ok (|d| >= EPS), L.  XP None: zero velocities."""
    X = np.asarray(X, np.float64)
    Vv = (np.zeros_like(X) if XP is None else (X - np.asarray(XP, np.float64)) / h)
    t, a, b, p0, p1 = _unpack(rows)
    m = len(t)
    bb = np.where(b >= 0, b, 0)
    xa, va = X[a, :3], Vv[a, :3]
    floor = t == FLOOR
    attach = t == ATTACH
    pb = np.where(attach[:, None], np.asarray(anchors, np.float64)[:, :3], X[bb, :3])
    d = xa - pb
    L = np.linalg.norm(d, axis=1)
    ok = np.where(floor, True, L >= EPS)
    n = np.where(floor[:, None], np.array([0.0, 1.0, 0.0]),
                 np.where(ok[:, None], d / np.where(L > 0, L, 1.0)[:, None], [0.0, 1.0, 0.0]))
    C = np.where(floor, xa[:, 1] - p0, L - p0)
    wb = np.where((t == DISTANCE) | (t == CONTACT), X[bb, 3], 0.0)
    w = X[a, 3] + wb
    alpha = np.where((t == DISTANCE) | (t == ATTACH), p1 / (h * h), 0.0)
    pen = np.where((t == FLOOR) | (t == CONTACT), np.maximum(0.0, -C), 0.0)
    vrel = np.where(((t == DISTANCE) | (t == CONTACT))[:, None], va - Vv[bb, :3], va)
    vn = np.sum(vrel * n, axis=1)
    vt = np.linalg.norm(vrel - n * vn[:, None], axis=1)
    return dict(t=t, a=a, b=b, C=C, w=w, alpha=alpha, n=n, vn=vn, vt=vt, pen=pen, ok=ok, L=L,
                m=m)


def features(X, XP, rows, anchors, h):
    """(m, 12) features in the kernel's order (slot 7 stays zero)."""
    g = geometry(X, XP, rows, anchors, h)
    s = g["w"] + g["alpha"]
    F = np.zeros((g["m"], 12))
    F[np.arange(g["m"]), g["t"]] = 1.0
    F[:, 4] = slog(g["C"] / LEN)
    F[:, 5] = np.log10(np.maximum(g["w"], 1e-30)) * 0.25
    F[:, 6] = np.where(s > 0, g["alpha"] / np.where(s > 0, s, 1.0), 0.0)
    F[:, 8] = slog(g["vn"] * h / LEN)
    F[:, 9] = slog(g["vt"] * h / LEN)
    F[:, 10] = slog(g["pen"] / LEN)
    F[:, 11] = np.log10(1.0 + g["alpha"]) * 0.125
    return F


def warm_launch(X, XP, rows, anchors, Y, h, relax=1.0):
    """The warm-start launch: every row with a first plain step lambda1 =
    -C / (w + alpha~) (one-sided rows only while C < 0) starts from
    lambda0 = 4 sigmoid(y) lambda1; each endpoint receives w_i n lambda0
    (-w_b for b) and a count of one; each particle with a count and w > 0
    moves by relax x the mean.  Returns the new positions (n, 4), lambda0
    (m,) and the (n, 4) sums/counts."""
    X = np.asarray(X, np.float64)
    g = geometry(X, XP, rows, anchors, h)
    s = g["w"] + g["alpha"]
    step = np.where((s > 0) & g["ok"], -g["C"] / np.where(s > 0, s, 1.0), 0.0)
    one_sided = (g["t"] == FLOOR) | (g["t"] == CONTACT)
    step = np.where(one_sided & (g["C"] >= 0), 0.0, step)
    lam0 = OMEGA_MAX / (1.0 + np.exp(-np.asarray(Y, np.float64))) * step
    acc = np.zeros((X.shape[0], 4))
    two = (g["t"] == DISTANCE) | (g["t"] == CONTACT)
    for j in np.nonzero(lam0 != 0.0)[0]:
        a, b = g["a"][j], g["b"][j]
        acc[a, :3] += g["n"][j] * (X[a, 3] * lam0[j])
        acc[a, 3] += 1.0
        if two[j]:
            acc[b, :3] += g["n"][j] * (-X[b, 3] * lam0[j])
            acc[b, 3] += 1.0
    out = X.copy()
    mv = (acc[:, 3] > 0) & (X[:, 3] > 0)
    out[mv, :3] += relax * acc[mv, :3] / acc[mv, 3][:, None]
    return out, lam0, acc


def residual_sum(X, rows, anchors, lam, h):
    """Sum over rows of r^2: r = C + alpha~ lambda (two-sided), max(0, -C)."""
    g = geometry(X, None, rows, anchors, h)
    one_sided = (g["t"] == FLOOR) | (g["t"] == CONTACT)
    r = np.where(one_sided, np.maximum(0.0, -g["C"]), g["C"] + g["alpha"] * np.asarray(lam, np.float64))
    return float(np.sum(r * r))
