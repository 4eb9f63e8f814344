"""Launch-shape autotuner for the 3-DOF Jacobi sweep.

For a flushed ``Solver`` in Jacobi mode it times whole frames (hipGraph
replays, HIP events on the launch stream) under each candidate launch shape
and keeps the fastest:

* ``tile``   — the fused per-iteration kernel with tile sizes of 32 ... 1024
  particles (44 B of LDS per particle; larger tiles re-project fewer halo
  rows, smaller tiles spread the launch over more CUs), each with integrate
  folded into the first iteration or run as its own launch, and with two
  iterations per substep also as two launches instead of one
  (``xpbd_jacobi_tile_pair`` re-projects the ring rows);
* ``window`` — projection with an LDS window + global atomics and a separate
  apply pass, with 1 or 2 rows per thread;
* ``block``  — whole frames in one workgroup (scenes that fit its LDS).

With a learned warm start, a per-particle gravity scale or device contact
generation (tile-mode extensions) only the tile candidates apply.

The simulation state is saved before and restored after each trial, so tuning
does not advance the scene.  Results are cached per scene signature (particle
count, row count, substeps, iterations) for the life of the process.
"""
from __future__ import annotations

_CACHE: dict = {}


def _signature(s):
    warm = s._ws_theta is not None
    return (s.num_bodies, s.num_constraints, s.substeps, s.iterations, s.solve_mode, warm,
            int(s.warm_start_every) if warm else 0, int(s.warm_start_judge) if warm else 0,
            s._gscale is not None, getattr(s, "_device_contacts", False))


def candidates(n_particles: int, tile_only: bool = False, fused_only: bool = False,
               pair: bool = False):
    """Launch shapes to time; ``tile_only`` (learned warm start): no window
    or block shapes; ``fused_only`` (the learned start with its judge, which
    runs in the fused plan): integrate folded in; ``pair`` (two plain
    iterations, or the learned start and one iteration): the fused plan also
    without the one-launch substep (fuse "two")."""
    out = []
    for t in (32, 64, 128, 256, 512, 1024):
        if t <= max(64, 2 * n_particles):
            out.append(("tile", t, None, True))
            if pair:
                out.append(("tile", t, None, "two"))
            if not fused_only:
                out.append(("tile", t, None, False))
    if tile_only:
        return out
    out.append(("window", 0, (1, None), False))
    out.append(("window", 0, (2, None), False))
    if n_particles <= _block_capacity():
        out.append(("block", 0, None, False))
    return out


def _block_capacity() -> int:
    from . import _native as N

    return int(N.lib().xpbd_frame_block_capacity(0))


def _apply(s, cand):
    kind, tile, tiling, fuse = cand
    s.block_frames = kind == "block"
    s.jacobi_kernel = "tile" if kind == "block" else kind
    s.tile_size = tile
    s.fuse_first = True if fuse == "two" else fuse
    s.tile_pair = False if fuse == "two" else "auto"
    s._dirty = True
    s._flush()
    if kind == "window" and tiling is not None:
        rpt, _ = tiling
        s._tiling = (rpt, s._tiling[1])


def tune(s, frames: int = 3, warmup: int = 1, verbose: bool = False):
    """Pick the fastest Jacobi launch shape for solver ``s`` (in place) and
    return ``(candidate, {candidate: ms_per_frame})``."""
    import torch

    if s.solve_mode != "jacobi":
        return None, {}
    s._flush()
    sig = _signature(s)
    if sig in _CACHE:
        _apply(s, _CACHE[sig])
        return _CACHE[sig], {}
    x0, v0 = s.positions(), s.velocities()
    times = {}
    tile_only = (s._ws_theta is not None or s._gscale is not None
                 or getattr(s, "_device_contacts", False))
    judged = s._ws_theta is not None and int(s.warm_start_judge) > 0 and s.substeps >= 2
    # plans that can run a substep as one launch: two plain iterations, or
    # the learned start and one iteration
    pair = ((s.iterations == 2 and s._ws_theta is None)
            or (s.iterations == 1 and s._ws_theta is not None))
    for cand in candidates(s.num_bodies, tile_only=tile_only, fused_only=judged, pair=pair):
        _apply(s, cand)
        s.set_state(x0, v0)
        s.step(warmup)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        s.step(frames)
        e1.record()
        torch.cuda.synchronize()
        times[cand] = e0.elapsed_time(e1) / frames
        if verbose:
            print(f"[autotune] {cand}: {times[cand]:.4f} ms/frame")
    best = min(times, key=times.get)
    _apply(s, best)
    s.set_state(x0, v0)
    _CACHE[sig] = best
    return best, times


# ---------------------------------------------------------------------------
# 6-DOF: leaf size of the LBVH broad phase
# ---------------------------------------------------------------------------
LEAF_SIZES = (1, 2, 4, 8, 16)
_ALL_PAIRS_MAX = 16384       # all-pairs candidate up to this many bodies


def leaf_candidates(n_bodies: int):
    """LBVH leaf sizes to time (0 = the all-pairs test)."""
    out = [0] if n_bodies <= _ALL_PAIRS_MAX else []
    return out + [L for L in LEAF_SIZES if L < max(2, n_bodies)]


def tune6(s, frames: int = 3, warmup: int = 1, verbose: bool = False):
    """Pick the LBVH leaf size (bodies per leaf, or all pairs) of a
    ``Solver6DOF`` with the LBVH broad phase by timing whole frames (HIP
    events; each frame's broad phase is part of its step) from the same
    state; sets ``s.lbvh_leaf`` and returns ``(leaf, {leaf: ms_per_frame})``.
    Every leaf size gives the same pair list, so only the time differs.
    Scenes stepped as whole-frame workgroups (their broad phase is inside
    that kernel) are left alone."""
    import torch

    if s.broadphase != "lbvh":
        return None, {}
    s._flush()
    n = s.num_bodies
    if n == 0 or s._use_block(n):
        return None, {}
    sig = ("6dof", n, s.num_joints, s.substeps, s.iterations)
    if sig in _CACHE:
        s.lbvh_leaf = _CACHE[sig]
        return _CACHE[sig], {}
    b = s._buf
    keep = {k: b[k].clone() for k in ("X", "Q", "V", "W")}
    times = {}
    for leaf in leaf_candidates(n):
        s.lbvh_leaf = leaf
        for k, t in keep.items():
            b[k].copy_(t)
        for _ in range(warmup):
            s.step()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(frames):
            s.step()
        e1.record()
        torch.cuda.synchronize()
        times[leaf] = e0.elapsed_time(e1) / frames
        if verbose:
            print(f"[autotune] lbvh leaf {leaf}: {times[leaf]:.4f} ms/frame")
    best = min(times, key=times.get)
    s.lbvh_leaf = best
    for k, t in keep.items():
        b[k].copy_(t)
    _CACHE[sig] = best
    return best, times
