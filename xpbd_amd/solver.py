"""3-DOF XPBD particle solver — the drop-in for ``xpbd3d.Solver``
(src/xpbd3d/solver.py:57-508), running on hand-written gfx950 HIP kernels.

Same constructor, scene-building methods, knobs and read-back as the
reference.  The host keeps the scene as growable columns and ``_flush``
packs them into device buffers (torch tensors used purely as HIP
memory) in the layout of include/xpbd.h:

* particles: ``X`` float4 (x, y, z, inv_mass), ``V``/``XP``/``ACC`` float4;
* rows: one int4 per constraint (body_a | type << 28, body_b, p0, p1),
  physically reordered — by color for Gauss-Seidel (each color is one
  contiguous launch), by smallest endpoint for Jacobi (so a tile's endpoints
  fall in one LDS window).  ``_perm`` maps device rows back to API order.

``step()`` runs one frame through ``xpbd_frame3_run``: the whole substep loop
is captured once into a hipGraph and replayed per frame (the reference
launches every kernel from Python).  There is no CPU execution path;
``device="cpu"`` is rejected.
"""
from __future__ import annotations

import numpy as np

from . import _native as N
from .coloring import color_constraints, color_ranges, color_summary
from .scene import Body, ConstraintHandle, Shape

# Constraint type codes (solver.py:39-43)
DISTANCE = 0
ATTACH = 1
FLOOR = 2
CONTACT = 3

MODE_GS = 0
MODE_JACOBI = 1
MODE_JACOBI_TILE = 2


def default_tile_size(n_particles: int) -> int:
    """Particles per tile of the fused Jacobi kernel: large tiles keep the
    share of re-projected halo rows small; small scenes get smaller tiles so
    that the launch still spreads over the 256 CUs.  512 is the largest
    default: on 8 M-particle cloths it beats 1024 (22 KB against 45 KB of LDS
    per block, so more blocks share a CU)."""
    for t in (512, 256, 128):
        if n_particles >= 256 * t:
            return t
    return 64


class _Col:
    """Growable host column: O(1) scalar appends, bulk numpy extends kept as
    chunks and joined once on demand.  Floats are held as float64 (the
    reference keeps Python floats in lists and rounds them to float32 at
    flush, solver.py:358-392)."""

    def __init__(self, dtype, width: int = 0):
        self.dtype = dtype
        self.width = width
        self._chunks: list = []
        self._n_chunks = 0
        self._tail: list = []

    def _shape(self):
        return (-1, self.width) if self.width else (-1,)

    def __len__(self):
        return self._n_chunks + len(self._tail)

    def append(self, v):
        self._tail.append(v)

    def extend(self, arr):
        self._spill()
        a = np.asarray(arr, self.dtype).reshape(self._shape())
        self._chunks.append(a)
        self._n_chunks += a.shape[0]

    def _spill(self):
        if self._tail:
            t = np.asarray(self._tail, self.dtype).reshape(self._shape())
            self._chunks.append(t)
            self._n_chunks += t.shape[0]
            self._tail = []

    def array(self) -> np.ndarray:
        self._spill()
        if len(self._chunks) != 1:
            empty = np.zeros((0, self.width) if self.width else (0,), self.dtype)
            self._chunks = [np.concatenate(self._chunks) if self._chunks else empty]
        return self._chunks[0]

    def set(self, arr) -> None:
        a = np.array(arr, self.dtype).reshape(self._shape())
        self._chunks = [a]
        self._n_chunks = a.shape[0]
        self._tail = []

    def __getitem__(self, i):
        if i >= self._n_chunks:
            return self._tail[i - self._n_chunks]
        return self.array()[i]


def _resolve_device(device) -> str:
    d = str(device)
    if d.startswith("hip"):
        d = "cuda" + d[3:]
    if not d.startswith("cuda"):
        raise ValueError(
            f"device {device!r}: this solver runs only on the HIP device path "
            "(use 'cuda:0' / 'hip:0'); there is no CPU implementation")
    return d


def pack_rows(types, a, b, rest, compliance, anchor, mu) -> tuple[np.ndarray, np.ndarray]:
    """Reference-layout row arrays -> (int32[M,4] packed rows, float32[M,4] anchors)."""
    m = len(types)
    types = np.asarray(types, np.int32)
    a = np.asarray(a, np.int32)
    b = np.asarray(b, np.int32)
    rest = np.asarray(rest, np.float32)
    compliance = np.asarray(compliance, np.float32)
    anchor = np.asarray(anchor, np.float32).reshape(-1, 3)
    mu = np.asarray(mu, np.float32)
    if m and int(a.max()) >= (1 << 28):
        raise ValueError("too many particles for the packed row format (2^28)")
    p0 = np.where(types == FLOOR, anchor[:, 1] if m else rest, rest).astype(np.float32)
    p1 = np.where((types == FLOOR) | (types == CONTACT), mu, compliance).astype(np.float32)
    rows = np.empty((m, 4), np.int32)
    rows[:, 0] = a | (types << 28)
    rows[:, 1] = b
    rows[:, 2] = p0.view(np.int32)
    rows[:, 3] = p1.view(np.int32)
    anc = np.zeros((m, 4), np.float32)
    anc[:, :3] = anchor
    return rows, anc


def jacobi_row_order(a, b) -> np.ndarray:
    """Stable order of rows by smallest endpoint (the LDS-window tiling key)."""
    a = np.asarray(a, np.int64)
    b = np.asarray(b, np.int64)
    key = np.where(b >= 0, np.minimum(a, b), a)
    return np.argsort(key, kind="stable")


# Row classes inside a tile of the fused kernel (the segments of
# xpbd_tile_compact): DISTANCE rows stored compactly, rows read as int4,
# and FLOOR rows of particles with exactly one FLOOR row (per-particle).
CLS_COMPACT, CLS_GENERAL, CLS_FLOOR = 0, 1, 2
COMPACT_SPAN = 1 << 22      # hi - tile base must stay below this (22-bit field)


def tile_row_class(a, b, types, tile: int, n_particles: int) -> np.ndarray:
    """Class of each row (CLS_*) for tiles of ``tile`` particles."""
    a = np.asarray(a, np.int64)
    b = np.asarray(b, np.int64)
    types = np.asarray(types, np.int64)
    lo = np.where(b >= 0, np.minimum(a, b), a)
    hi = np.where(b >= 0, np.maximum(a, b), a)
    cls = np.full(len(a), CLS_GENERAL, np.int64)
    if tile <= 1024:
        cls[(types == DISTANCE) & (hi - (lo // tile) * tile < COMPACT_SPAN)] = CLS_COMPACT
    fl = types == FLOOR
    if np.any(fl):
        per = np.bincount(a[fl], minlength=max(n_particles, 1))
        cls[fl & (per[np.where(fl, a, 0)] == 1)] = CLS_FLOOR
    return cls


def jacobi_tile_order(a, b, types, tile: int, n_particles: int | None = None) -> np.ndarray:
    """Stable row order of the fused tile kernel: grouped by the tile of the
    smallest endpoint; inside a tile by class (compact DISTANCE rows, rows
    read whole, single FLOOR rows: the segments of the compact tables), then
    by type, then by rank among the rows of the same smallest endpoint, then
    by that endpoint.  Rows of a wavefront then take the same branch and
    mostly add into distinct particles."""
    a = np.asarray(a, np.int64)
    b = np.asarray(b, np.int64)
    lo = np.where(b >= 0, np.minimum(a, b), a)
    if n_particles is None:
        n_particles = int(max(a.max(initial=-1), b.max(initial=-1))) + 1
    cls = tile_row_class(a, b, types, tile, n_particles)
    first = jacobi_row_order(a, b)
    los = lo[first]
    rank = np.empty(len(lo), np.int64)
    rank[first] = np.arange(len(lo), dtype=np.int64) - np.searchsorted(los, los, "left")
    rank[cls == CLS_FLOOR] = 0
    key = np.lexsort((lo, rank, np.asarray(types, np.int64), cls, lo // tile))
    return key.astype(np.int64)


def compact_tile_tables(rows_sorted: np.ndarray, n_particles: int, tile: int, tile_row_start,
                        halo_rows) -> dict:
    """Compact tables (xpbd_tile_compact) of rows in ``jacobi_tile_order``:
    per tile the ends of the compact and the general segment, the compact
    rows (ci int32[M, 2]: tile-relative (lo, hi) | rest bits; cc float32[M]
    compliance), the halo entries (int32[H, 2]: row, owner tile base or -1)
    and the per-particle floors (pf float32[n, 2] floor_y / mu, floor_y =
    -inf without one; pf_rows / pf_part: each such FLOOR row and its
    particle, to read their multipliers back), and the counts of rows read
    whole (n_general) and of per-particle FLOOR rows (n_floor)."""
    m = rows_sorted.shape[0]
    n_tiles = len(tile_row_start) - 1
    t = (rows_sorted[:, 0] >> 28).astype(np.int64)
    a = (rows_sorted[:, 0] & 0x0FFFFFFF).astype(np.int64)
    b = rows_sorted[:, 1].astype(np.int64)
    cls = tile_row_class(a, b, t, tile, n_particles)
    lo = np.where(b >= 0, np.minimum(a, b), a)
    hi = np.where(b >= 0, np.maximum(a, b), a)
    lt = lo // tile
    base = lt * tile
    if m and np.any(np.diff(lt * 4 + cls) < 0):
        raise ValueError("rows are not in jacobi_tile_order")
    cnt = np.zeros((n_tiles, 3), np.int64)
    np.add.at(cnt, (lt, cls), 1)
    trs = np.asarray(tile_row_start, np.int64)
    dist_end = (trs[:-1] + cnt[:, 0]).astype(np.int32)
    gen_end = (dist_end + cnt[:, 1]).astype(np.int32)
    ci = np.zeros((m, 2), np.int32)
    c = cls == CLS_COMPACT
    ci[c, 0] = ((lo[c] - base[c]) | ((hi[c] - base[c]) << 10)).astype(np.uint32).view(np.int32)
    ci[:, 1] = rows_sorted[:, 2]
    cc = np.ascontiguousarray(rows_sorted[:, 3]).view(np.float32).copy()
    hr = np.asarray(halo_rows, np.int64)
    halo2 = np.zeros((len(hr), 2), np.int32)
    halo2[:, 0] = hr
    halo2[:, 1] = np.where(cls[hr] == CLS_COMPACT, base[hr], -1)
    pf = np.zeros((max(n_particles, 1), 2), np.float32)
    pf[:, 0] = -np.inf
    pf_rows = np.nonzero(cls == CLS_FLOOR)[0]
    pf_part = a[pf_rows]
    pf[pf_part, 0] = rows_sorted[pf_rows, 2].view(np.float32)
    pf[pf_part, 1] = rows_sorted[pf_rows, 3].view(np.float32)
    return dict(dist_end=dist_end, gen_end=gen_end, ci=ci, cc=cc, halo2=halo2, pf=pf,
                pf_rows=pf_rows, pf_part=pf_part,
                n_general=int(np.count_nonzero(cls == CLS_GENERAL)), n_floor=len(pf_rows))


def choose_jacobi_tiling(rows_sorted: np.ndarray, block: int = 256) -> tuple[int, int]:
    """(rows_per_thread, window) for the LDS-staged Jacobi projection: the
    smallest power-of-two window that holds the endpoint span of 90 % of the
    tiles (the others spill to global atomics)."""
    m = rows_sorted.shape[0]
    if m == 0:
        return 1, 0
    rpt = 2 if m >= 64 * 1024 else 1
    tile = block * rpt
    a = rows_sorted[:, 0] & 0x0FFFFFFF
    b = rows_sorted[:, 1]
    lo = np.where(b >= 0, np.minimum(a, b), a)
    hi = np.maximum(a, b)
    nt = (m + tile - 1) // tile
    pad = nt * tile - m
    lo_t = np.pad(lo, (0, pad), mode="edge").reshape(nt, tile)[:, 0]
    hi_t = np.pad(hi, (0, pad), mode="edge").reshape(nt, tile).max(axis=1)
    span = (hi_t - lo_t + 1).astype(np.int64)
    want = int(np.percentile(span, 90))
    window = 256
    while window < want and window < 4096:
        window *= 2
    return rpt, window


def morton_order(pos: np.ndarray) -> np.ndarray:
    """Particle order along a 3-D Z-order curve (21 bits per axis over the
    bounding cube of ``pos``): a permutation that keeps spatial neighbours,
    and therefore the endpoints of a constraint, close in memory.  One scale
    for all axes: a flat sheet with 1e-3 of jitter across it keeps that
    axis in the low bits (scaled per axis, the jitter would fill all 21 bits
    and order the sheet by noise: 67 % of a 128^2 cloth's rows would cross
    64-particle tiles instead of 16 %)."""
    n = pos.shape[0]
    if n < 2:
        return np.arange(n, dtype=np.int64)
    lo = pos.min(axis=0)
    ext = max(float((pos.max(axis=0) - lo).max()), 1e-30)
    q = np.clip(((pos - lo) / ext * (1 << 21)).astype(np.int64), 0, (1 << 21) - 1)
    code = np.zeros(n, dtype=np.uint64)
    for bit in range(21):
        for ax in range(3):
            code |= ((q[:, ax] >> bit) & 1).astype(np.uint64) << np.uint64(3 * bit + 2 - ax)
    return np.argsort(code, kind="stable")


def tile_tables(rows_sorted: np.ndarray, n_particles: int, tile: int):
    """Tile tables of the fused Jacobi kernel for rows grouped by the tile of
    their smallest endpoint: ``tile_row_start`` (first owned row of each
    tile, n_tiles + 1 entries) and the halo rows (larger endpoint in a later
    tile than the smaller one) grouped by the tile of the larger endpoint."""
    n_tiles = max(1, (n_particles + tile - 1) // tile)
    a = (rows_sorted[:, 0] & 0x0FFFFFFF).astype(np.int64)
    b = rows_sorted[:, 1].astype(np.int64)
    lo = np.where(b >= 0, np.minimum(a, b), a)
    hi = np.where(b >= 0, np.maximum(a, b), a)
    lt = lo // tile
    if len(lt) and np.any(np.diff(lt) < 0):
        raise ValueError("rows are not grouped by the tile of their smallest endpoint")
    starts = np.searchsorted(lt, np.arange(n_tiles + 1, dtype=np.int64), "left")
    starts[-1] = len(lo)
    cross = np.nonzero((hi // tile) != (lo // tile))[0]
    ht = hi[cross] // tile
    order = np.argsort(ht, kind="stable")
    halo_rows = cross[order].astype(np.int32)
    halo_start = np.searchsorted(ht[order], np.arange(n_tiles + 1), "left").astype(np.int32)
    return starts.astype(np.int32), halo_start, halo_rows


def _ranges(starts, lens) -> np.ndarray:
    """Concatenation of arange(s, s + l) for each (s, l)."""
    starts = np.asarray(starts, np.int64)
    lens = np.asarray(lens, np.int64)
    total = int(lens.sum())
    if total == 0:
        return np.zeros(0, np.int64)
    off = np.repeat(np.cumsum(lens) - lens, lens)
    return np.repeat(starts, lens) + (np.arange(total, dtype=np.int64) - off)


def tile_pair_tables(rows_sorted: np.ndarray, n_particles: int, tile: int):
    """Tables of the two-iteration tile launch (xpbd_jacobi_tile_pair) for
    rows in the tile kernel's order.  Per tile: ring 1 (particles one row
    away from the tile) and ring 2 (one row away from ring 1), the phase-1
    rows (every row incident to the tile or ring 1, with local endpoint
    indices: tile particle p -> p - p0, ring k -> tile + its position) and
    the phase-2 rows (those incident to the tile, as positions in the
    phase-1 list, bit 31 set where the tile owns the row: its smallest
    endpoint lies in the tile).  Returns a dict of int32 arrays and the
    maxima (local particles, tile + ring 1, phase-1 rows) that size the
    launch's LDS."""
    m = rows_sorted.shape[0]
    n_tiles = max(1, (n_particles + tile - 1) // tile)
    a = (rows_sorted[:, 0] & 0x0FFFFFFF).astype(np.int64)
    b = rows_sorted[:, 1].astype(np.int64)
    two = b >= 0
    ends = np.concatenate([a, b[two]])
    rid = np.concatenate([np.arange(m, dtype=np.int64), np.nonzero(two)[0]])
    o = np.argsort(ends, kind="stable")
    inc_rows = rid[o]
    inc_start = np.searchsorted(ends[o], np.arange(n_particles + 1, dtype=np.int64), "left")
    lo = np.where(two, np.minimum(a, b), a)

    def incident(parts):
        rr = inc_rows[_ranges(inc_start[parts], inc_start[parts + 1] - inc_start[parts])]
        return np.unique(rr)

    def endpoints(rr):
        e = np.concatenate([a[rr], b[rr][two[rr]]])
        return np.unique(e)

    ring_start, ring1_n, ring_idx = [0], [], []
    p1_start, p1, p2_start, p2 = [0], [], [0], []
    lmax = amax = k1max = 0
    for t in range(n_tiles):
        p0, p1e = t * tile, min((t + 1) * tile, n_particles)
        T = np.arange(p0, p1e, dtype=np.int64)
        rows_t = np.unique(inc_rows[inc_start[p0]:inc_start[p1e]])
        r1 = endpoints(rows_t)
        r1 = r1[(r1 < p0) | (r1 >= p1e)]
        rows_r1 = incident(r1)
        rows_1 = np.union1d(rows_t, rows_r1)
        r2 = endpoints(rows_r1)
        r2 = r2[((r2 < p0) | (r2 >= p1e)) & ~np.isin(r2, r1)]
        loc_g = np.concatenate([T, r1, r2])
        loc_l = np.concatenate([T - p0, tile + np.arange(len(r1) + len(r2), dtype=np.int64)])
        so = np.argsort(loc_g)
        sg, sl = loc_g[so], loc_l[so]

        def local(g):
            return sl[np.searchsorted(sg, g)]

        la = local(a[rows_1])
        lb = np.where(two[rows_1], local(np.where(two[rows_1], b[rows_1], p0)), -1)
        ent = np.zeros((len(rows_1), 4), np.int32)
        ent[:, 0], ent[:, 1], ent[:, 2] = rows_1, la, lb
        pos_t = np.searchsorted(rows_1, rows_t)
        own = (lo[rows_t] // tile) == t
        e2 = pos_t.astype(np.int64) | (own.astype(np.int64) << 31)
        ring_idx.append(np.concatenate([r1, r2]))
        ring1_n.append(len(r1))
        ring_start.append(ring_start[-1] + len(r1) + len(r2))
        p1.append(ent)
        p1_start.append(p1_start[-1] + len(rows_1))
        p2.append(e2)
        p2_start.append(p2_start[-1] + len(rows_t))
        lmax = max(lmax, tile + len(r1) + len(r2))
        amax = max(amax, tile + len(r1))
        k1max = max(k1max, len(rows_1))
    cat = lambda xs, dt: (np.concatenate(xs) if xs else np.zeros(0, dt))  # noqa: E731
    amax += amax & 1          # the 8-byte sums of amax particles precede 16-byte records
    lmax = max(lmax, amax)
    return {
        "ring_start": np.asarray(ring_start, np.int32),
        "ring1_n": np.asarray(ring1_n, np.int32),
        "ring_idx": cat(ring_idx, np.int64).astype(np.int32),
        "p1_start": np.asarray(p1_start, np.int32),
        "p1": cat(p1, np.int32).reshape(-1, 4).astype(np.int32),
        "p2_start": np.asarray(p2_start, np.int32),
        "p2": cat(p2, np.int64).astype(np.uint32).view(np.int32),
        "lmax": lmax, "amax": amax, "k1max": k1max,
    }


class Solver:
    """3D XPBD particle solver (solver.py:57-126).

    Args mirror the reference: ``dt`` frame step, ``substeps``,
    ``iterations`` per substep, ``gravity``, per-substep ``damping``,
    ``solve_mode`` ("gs", "jacobi" or "auto" = jacobi), ``jacobi_relax``,
    ``device`` ("cuda:N"/"hip:N"), ``max_speed`` (0 disables the clamp).
    """

    def __init__(
        self,
        dt: float = 1.0 / 60.0,
        substeps: int = 10,
        iterations: int = 1,
        gravity: tuple = (0.0, -9.81, 0.0),
        damping: float = 0.0,
        solve_mode: str = "auto",
        jacobi_relax: float = 1.0,
        device: str = "cuda:0",
        max_speed: float = 40.0,
    ):
        self.device = _resolve_device(device)
        self.dt = float(dt)
        self.substeps = int(substeps)
        self.iterations = int(iterations)
        self.gravity = tuple(float(g) for g in gravity)
        self.damping = float(damping)
        self.jacobi_relax = float(jacobi_relax)
        self.max_speed = float(max_speed)
        if solve_mode == "auto":
            solve_mode = "jacobi"
        assert solve_mode in ("gs", "jacobi")
        self.solve_mode = solve_mode
        self.use_graph = True
        # Jacobi kernel: "tile" (one fused launch per iteration) or "window"
        # (projection with LDS-window atomics + a separate apply pass)
        self.jacobi_kernel = "tile"
        self.tile_size = 0            # 0 = chosen from the scene size
        self.reorder = True           # Z-order particle layout for Jacobi
        # integrate folded into the tile launches: True = into the previous
        # substep's last launch where the rows allow it (no CONTACT rows),
        # else into the substep's first launch; "first" = always the latter;
        # False = a separate integrate pass (identical results)
        self.fuse_first = True
        # both iterations of a substep in one launch (xpbd_jacobi_tile_pair;
        # tile mode, 2 iterations): True, False or "auto" (scenes up to
        # PAIR_AUTO_ROWS rows, where a launch costs latency, not bytes)
        self.tile_pair = "auto"
        self.block_frames = "auto"    # small scenes: whole frames in one workgroup
        self._block_active = False
        # tile launches over the compact tables (xpbd_tile_compact): True,
        # or False for the int4 rows and halo list alone (identical results)
        self.compact_rows = True
        self._cx = self._cx_dev = self._cx_struct = None
        self._plan_pf = False         # the plan's FLOOR multipliers are per particle
        self._plan_pair = False       # the plan runs two-iteration pair launches
        # device-side CONTACT generation (Jacobi tile mode): an extension
        # that replaces the per-frame host broad phase; see enable_device_contacts
        self._device_contacts = False
        self._gscale = None           # per-particle gravity factor (API order)
        self._ws_theta = None         # learned lambda warm start parameters
        self._ws_gen = 0
        self.warm_start_every = 1     # network pass every k-th substep of a frame
        # the zero start judges the learned one in the frame plan every k-th
        # frame (True / 1: every frame; False / 0: never).  Every frame costs
        # one extra substep per frame (8 % of the batched world's frame, 15 %
        # of the 128^2 cloth's); every 8th frame about 1 %
        self.warm_start_judge = 8
        self.WSON = self.WSST = None

        # host scene, one column per field (float64 as the reference's
        # Python floats; rounded to float32 once, at flush)
        self._x = _Col(np.float64, 3)
        self._v = _Col(np.float64, 3)
        self._mass = _Col(np.float64)
        self._radius = _Col(np.float64)
        self._collide = _Col(np.bool_)
        self._friction = _Col(np.float64)
        self._rt = _Col(np.int32)          # row type
        self._ra = _Col(np.int32)          # body_a
        self._rb = _Col(np.int32)          # body_b (-1 = none)
        # row scalars are rounded to float32 on entry: the reference stores
        # Python floats and rounds each once at flush, which is the same value
        self._rrest = _Col(np.float32)
        self._rcomp = _Col(np.float32)
        self._ranch = _Col(np.float32, 3)
        self._rmu = _Col(np.float32)

        self._self_collision = False
        self._contact_margin = 0.0
        self._default_friction = 0.0
        self._contact_pool_start = None
        self._dirty = True

        # device buffers (torch tensors), built in _flush
        self.X = self.V = self.XP = self.ACC = None
        self.R = self.ANC = self.LAM = None
        self._perm = np.zeros(0, np.int64)
        self._color_starts = np.zeros(1, dtype=np.int32)
        self.num_colors = 0
        self.color_counts: dict = {}
        self._tiling = (1, 0)
        self._pperm = None            # device particle k <- API particle _pperm[k]
        self._pinv = None             # API particle i -> device particle _pinv[i]
        self.X_B = self.LAM_B = None
        self._tile_tabs = None
        self._plan = None
        self._plan_sig = None
        self._gen = 0

    # ---- scene building (solver.py:130-228) ------------------------------
    def add_particle(self, position, mass, velocity=(0.0, 0.0, 0.0), shape=None, collide=True,
                     radius=None, friction=None) -> Body:
        """One particle (solver.py:130-151).  Its collision radius is
        ``radius`` or the shape's bounding radius; it collides only when asked
        to and when it has both a radius and a mass; friction is clamped at 0."""
        sh = Shape() if shape is None else shape
        r = _shape_radius(sh) if radius is None else float(radius)
        m = float(mass)
        mu = self._default_friction if friction is None else float(friction)
        handle = Body(index=len(self._x), shape=sh)
        for column, value in ((self._x, tuple(map(float, position))),
                              (self._v, tuple(map(float, velocity))),
                              (self._mass, m), (self._radius, r),
                              (self._collide, bool(collide) and r > 0.0 and m > 0.0),
                              (self._friction, max(0.0, mu))):
            column.append(value)
        self._dirty = True
        return handle

    def add_particles(self, positions, masses, velocities=None, radius=0.0, collide=False,
                      friction=None) -> np.ndarray:
        """Bulk ``add_particle`` (an extension for million-particle scenes):
        appends len(positions) particles, returns their indices."""
        pos = np.asarray(positions, np.float64).reshape(-1, 3)
        n = pos.shape[0]
        m = np.broadcast_to(np.asarray(masses, np.float64), (n,))
        vel = (np.zeros((n, 3)) if velocities is None
               else np.asarray(velocities, np.float64).reshape(-1, 3))
        r = np.broadcast_to(np.asarray(radius, np.float64), (n,))
        mu = self._default_friction if friction is None else friction
        mu = np.maximum(0.0, np.broadcast_to(np.asarray(mu, np.float64), (n,)))
        base = len(self._x)
        self._x.extend(pos)
        self._v.extend(vel)
        self._mass.extend(m)
        self._radius.extend(r)
        self._collide.extend(np.asarray(collide, bool) & (r > 0.0) & (m > 0.0))
        self._friction.extend(mu)
        self._dirty = True
        return np.arange(base, base + n, dtype=np.int64)

    def _extend_rows(self, t, a, b, rest=0.0, comp=0.0, anchor=None, mu=0.0) -> np.ndarray:
        """Append rows of type ``t`` over index arrays (scalars broadcast);
        returns their row indices."""
        a = np.asarray(a, np.int64).ravel()
        n = a.shape[0]
        base = len(self._rt)
        self._rt.extend(np.full(n, t, np.int32))
        self._ra.extend(a)
        self._rb.extend(np.asarray(b, np.int64).ravel())
        self._rrest.extend(np.broadcast_to(np.asarray(rest, np.float64), (n,)))
        self._rcomp.extend(np.broadcast_to(np.asarray(comp, np.float64), (n,)))
        self._ranch.extend(np.zeros((n, 3)) if anchor is None else np.asarray(anchor))
        self._rmu.extend(np.broadcast_to(np.asarray(mu, np.float64), (n,)))
        self._dirty = True
        return np.arange(base, base + n, dtype=np.int64)

    def _add_row(self, t, a, b, rest=0.0, comp=0.0, anchor=(0.0, 0.0, 0.0), mu=0.0) -> int:
        idx = len(self._rt)
        self._rt.append(t)
        self._ra.append(a)
        self._rb.append(b)
        self._rrest.append(rest)
        self._rcomp.append(comp)
        self._ranch.append(anchor)
        self._rmu.append(mu)
        self._dirty = True
        return idx

    def add_distance(self, body_a, body_b, rest=None, compliance=0.0) -> ConstraintHandle:
        """Compliant distance ``|x_a - x_b| = rest`` (rest None = current separation)."""
        if rest is None:
            pa = np.asarray(self._x[body_a.index], np.float64)
            pb = np.asarray(self._x[body_b.index], np.float64)
            rest = float(np.linalg.norm(pa - pb))
        idx = self._add_row(DISTANCE, body_a.index, body_b.index, float(rest), float(compliance))
        return ConstraintHandle(index=idx)

    def add_distances(self, a, b, rest=None, compliance=0.0) -> np.ndarray:
        """Bulk ``add_distance`` over index arrays (an extension); ``rest=None``
        takes each pair's current separation."""
        a = np.asarray(a, np.int64).ravel()
        b = np.asarray(b, np.int64).ravel()
        n = a.shape[0]
        if rest is None:
            x = self._x.array()
            rest = np.linalg.norm(x[a] - x[b], axis=1)
        return self._extend_rows(DISTANCE, a, b, rest=rest, comp=compliance)

    def add_attach(self, body, world_point, compliance=0.0, rest=0.0) -> ConstraintHandle:
        """Compliant pin of a particle to a world point."""
        idx = self._add_row(ATTACH, body.index, -1, float(rest), float(compliance),
                            tuple(float(p) for p in world_point))
        return ConstraintHandle(index=idx)

    def add_floor_contact(self, body, floor_y=0.0, friction=None) -> ConstraintHandle:
        """One-sided floor at ``y = floor_y`` with Coulomb friction."""
        mu = float(friction) if friction is not None else float(self._friction[body.index])
        idx = self._add_row(FLOOR, body.index, -1, anchor=(0.0, float(floor_y), 0.0),
                            mu=max(0.0, mu))
        return ConstraintHandle(index=idx)

    def add_floor_contacts(self, bodies, floor_y=0.0, friction=None) -> np.ndarray:
        """Bulk ``add_floor_contact`` over an index array (an extension)."""
        a = np.asarray(bodies, np.int64).ravel()
        n = a.shape[0]
        mu = (self._friction.array()[a] if friction is None
              else np.broadcast_to(np.asarray(friction, np.float64), (n,)))
        anch = np.zeros((n, 3))
        anch[:, 1] = float(floor_y)
        return self._extend_rows(FLOOR, a, np.full(n, -1), anchor=anch, mu=np.maximum(0.0, mu))

    def enable_self_collision(self, enabled=True, margin=0.0, default_friction=None) -> None:
        """Per-frame sphere-sphere CONTACT generation between ``collide`` bodies."""
        self._self_collision = bool(enabled)
        self._contact_margin = float(margin)
        if default_friction is not None:
            self._default_friction = max(0.0, float(default_friction))
            fr = self._friction.array().copy()
            fr[fr == 0.0] = self._default_friction
            self._friction.set(fr)
            self._refresh_contact_friction()
        if enabled and self._contact_pool_start is None:
            self._contact_pool_start = len(self._rt)

    def set_gravity_scale(self, scale) -> None:
        """Per-particle gravity factor g_i = g * scale[i] (an extension used
        for batches of independent scenes; Jacobi tile mode only)."""
        self._gscale = np.asarray(scale, np.float32).reshape(-1).copy()
        self._dirty = True

    def set_warm_start(self, theta) -> None:
        """Learned lambda warm start (an extension; the reference resets
        lambda to zero every substep, solver.py:431): ``theta`` holds the
        5057 parameters of the 12-64-64-1 network (``warmstart.train``);
        every substep then starts each row from the predicted multiplier
        and the matching position offset (``xpbd_ws_net`` +
        ``xpbd_jacobi_tile_warm``) before its Jacobi iterations.  ``None``
        restores the zero start.
        Runs the Jacobi tile kernel.  The plain zero start stays the judge
        and fallback (``warm_start_judge`` = k, every 8th frame by default): the
        first substep of every k-th frame is also run with the other start
        from the same state, and the learned start stays on for the rest of
        the frame and the following ones only while its residual is no
        larger (``warm_start_stats``)."""
        import torch

        if theta is None:
            self._ws_theta = None
        else:
            if self.solve_mode != "jacobi":
                raise ValueError("the learned warm start runs with solve_mode='jacobi'")
            t = torch.as_tensor(theta, dtype=torch.float32).to(self.device).reshape(-1).clone()
            n = int(N.lib().xpbd_mlp_param_count())
            if t.numel() != n:
                raise ValueError(f"warm-start parameters: {t.numel()} values, expected {n}")
            self._ws_theta = t
            self.jacobi_kernel = "tile"
        self._ws_gen += 1
        self._dirty = True

    def enable_device_contacts(self, enabled=True, margin=0.0) -> None:
        """Sphere-sphere CONTACT rows generated on the GPU every frame (an
        extension; the reference rebuilds them on the host, solver.py:275-342).
        Same margin, cell and filter as the host path, same rows in the same
        order; pairs already joined by a two-body constraint are skipped, as
        the host path skips them (solver.py:285-290, 336-337).  Jacobi: the
        tile kernel projects them each iteration.  Gauss-Seidel: every frame
        the rows (static and contact) are recoloured on the device with the
        reference's greedy colouring (``xpbd_color_rows_device``) and the
        whole frame runs in one workgroup (``xpbd_frame_block_ordered``),
        for scenes of up to 4096 particles."""
        if enabled and self.solve_mode == "jacobi" and self.jacobi_kernel != "tile":
            raise ValueError("device contacts in Jacobi mode need the tile kernel")
        self._device_contacts = bool(enabled)
        self._contact_margin = float(margin)
        self._dirty = True

    # ---- runtime perturbations (solver.py:232-271); in place on device ----
    def _dev(self, i: int) -> int:
        return int(i) if self._pinv is None else int(self._pinv[int(i)])

    def set_position(self, body, p) -> None:
        self._flush()
        self.X[self._dev(body.index), :3] = self._t(np.asarray(p, np.float32))

    def set_velocity(self, body, vel) -> None:
        self._flush()
        self.V[self._dev(body.index), :3] = self._t(np.asarray(vel, np.float32))

    def add_impulse(self, body, dv) -> None:
        self._flush()
        self.V[self._dev(body.index), :3] += self._t(np.asarray(dv, np.float32))

    def set_state(self, x, v) -> None:
        """Overwrite all positions and velocities ((n, 3) arrays in API
        particle order) on the device."""
        self._flush()
        x = np.asarray(x, np.float32).reshape(-1, 3)
        v = np.asarray(v, np.float32).reshape(-1, 3)
        if self._pperm is not None:
            x, v = x[self._pperm], v[self._pperm]
        self.X[:, :3] = self._t(x)
        self.V[:, :3] = self._t(v)

    def set_all_compliance(self, value: float, only_distance: bool = True) -> None:
        """Live compliance update; rewrites the device rows in place (no
        reallocation, so a captured graph stays valid)."""
        self._flush()
        comp = self._rcomp.array().copy()
        sel = (self._rt.array() == DISTANCE) if only_distance else slice(None)
        comp[sel] = float(value)
        self._rcomp.set(comp)
        self._repack_rows_inplace()

    def set_all_friction(self, mu: float) -> None:
        """Live friction update (solver.py:261-271): every particle's
        friction, the FLOOR/CONTACT rows and the per-particle friction the
        device CONTACT rows are generated from, all rewritten in place."""
        self._flush()
        self._default_friction = float(mu)
        self._friction.set(np.full(len(self._friction), float(mu)))
        rmu = self._rmu.array().copy()
        t = self._rt.array()
        rmu[(t == FLOOR) | (t == CONTACT)] = float(mu)
        self._rmu.set(rmu)
        self._repack_rows_inplace()
        self._refresh_contact_friction()

    def _refresh_contact_friction(self) -> None:
        """Per-particle friction of the device CONTACT generator (device
        order), rewritten in place so a captured plan stays valid."""
        if getattr(self, "CMU", None) is None or self._dirty:
            return
        fr = self._friction.array().astype(np.float32)
        if self._pperm is not None:
            fr = fr[self._pperm]
        self.CMU.copy_(self._t(np.ascontiguousarray(fr)))

    # ---- host broad phase (solver.py:275-342) ------------------------------
    def _rebuild_contacts(self) -> None:
        """This frame's CONTACT rows (solver.py:275-342): last frame's are
        dropped, then every pair of colliding bodies within reach gets one,
        in the candidate order of ``grid_candidate_pairs``, unless a
        constraint already joins the pair.  Reach is the sum of the radii
        plus a margin that grows with the fastest velocity component
        (margin + 3 dt max|v|, at most the largest radius); the pair test
        and the row fields are float32 as on the device."""
        if not self._self_collision:
            return
        static = self._rt.array() != CONTACT
        for col in (self._rt, self._ra, self._rb, self._rrest, self._rcomp, self._ranch,
                    self._rmu):
            col.set(col.array()[static])
        self._contact_pool_start = len(self._rt)
        self._dirty = True
        bodies = np.flatnonzero(self._collide.array().astype(bool))
        if bodies.size < 2:
            return
        x = self.positions() if self.X is not None else self._x.array().astype(np.float32)
        rad = self._radius.array().astype(np.float32)
        fr = self._friction.array().astype(np.float32)
        fastest = float(np.abs(self.velocities()).max()) if self.V is not None else 0.0
        r_big = float(rad[bodies].max())
        margin = min(self._contact_margin + 3.0 * self.dt * fastest, r_big)
        la, lb = grid_candidate_pairs(x[bodies], r_big * 2.0 + margin)
        if la.size == 0:
            return
        ia, ib = bodies[la], bodies[lb]
        gap = x[ia] - x[ib]
        reach = rad[ia] + rad[ib] + margin
        near = np.einsum("ij,ij->i", gap, gap) <= reach * reach
        ia, ib = ia[near], ib[near]
        # drop pairs a constraint joins: keys lo << 32 | hi of both sets
        ra, rb = self._ra.array(), self._rb.array()
        two = rb >= 0
        joined = np.unique((np.minimum(ra[two], rb[two]).astype(np.int64) << 32)
                           | np.maximum(ra[two], rb[two]).astype(np.int64))
        key = (np.minimum(ia, ib) << 32) | np.maximum(ia, ib)
        free = ~np.isin(key, joined, assume_unique=False)
        ia, ib = ia[free], ib[free]
        if ia.size:
            self._extend_rows(CONTACT, ia, ib, rest=rad[ia] + rad[ib],
                              mu=np.sqrt(fr[ia] * fr[ib]))

    # ---- flush: host lists -> device buffers (solver.py:346-408) ----------
    def host_arrays(self) -> dict:
        """The scene in the reference's own array layout (solver.py:358-392):
        float32 x/v (n,3), inv_mass (n,), row type/a/b (int32), rest,
        compliance, mu (float32) and anchor (m,3)."""
        x_np = self._x.array().astype(np.float32)
        v_np = self._v.array().astype(np.float32)
        mass = self._mass.array().astype(np.float32)
        inv_mass = np.where(mass > 0.0, 1.0 / np.maximum(mass, 1e-20), 0.0).astype(np.float32)
        return dict(
            x=x_np, v=v_np, inv_mass=inv_mass,
            type=self._rt.array().astype(np.int32),
            a=self._ra.array().astype(np.int32),
            b=self._rb.array().astype(np.int32),
            rest=self._rrest.array().astype(np.float32),
            compliance=self._rcomp.array().astype(np.float32),
            anchor=self._ranch.array().astype(np.float32),
            mu=self._rmu.array().astype(np.float32),
        )

    def _t(self, arr):
        import torch

        return torch.as_tensor(arr).to(self.device)

    def _flush(self) -> None:
        if not self._dirty:
            return
        import torch

        N.lib()
        h = self.host_arrays()
        n_b = h["x"].shape[0]
        n_c = h["type"].shape[0]
        x_np, v_np = h["x"], h["v"]
        if self.X is not None:
            prev = self.X.shape[0]
            if 0 < prev <= n_b:
                x_np[:prev] = self.positions()
                v_np[:prev] = self.velocities()
        tile_mode = self.solve_mode == "jacobi" and self.jacobi_kernel == "tile"
        if tile_mode and self.reorder and n_b > 1:
            pperm = morton_order(x_np.astype(np.float64))
            pinv = np.empty_like(pperm)
            pinv[pperm] = np.arange(n_b, dtype=np.int64)
        else:
            pperm = pinv = None
        inv_mass = h["inv_mass"]
        ra, rb = h["a"], h["b"]
        if pperm is not None:
            x_np, v_np, inv_mass = x_np[pperm], v_np[pperm], inv_mass[pperm]
            ra = pinv[ra].astype(np.int32)
            rb = np.where(rb >= 0, pinv[np.maximum(rb, 0)], -1).astype(np.int32)
        X = np.zeros((n_b, 4), np.float32)
        X[:, :3] = x_np
        X[:, 3] = inv_mass
        V = np.zeros((n_b, 4), np.float32)
        V[:, :3] = v_np
        rows, anc = pack_rows(h["type"], ra, rb, h["rest"], h["compliance"],
                              h["anchor"], h["mu"])
        gs_dev = self.solve_mode == "gs" and self._device_contacts
        if gs_dev:
            cap = min(int(N.lib().xpbd_frame_block_capacity(1)),
                      int(N.lib().xpbd_color_rows_max_bodies()))
            if n_b > cap:
                raise ValueError(f"Gauss-Seidel with device contacts runs scenes of up to {cap} "
                                 f"particles in one workgroup ({n_b} given)")
            perm = np.arange(n_c, dtype=np.int64)      # rows stay in API order
            self._color_starts = np.array([0, n_c], dtype=np.int32)
            self.num_colors = 0
            self.color_counts = {}
        elif self.solve_mode == "gs" and n_c > 0:
            color = color_constraints(n_b, h["a"], h["b"])
            order, starts, num_colors = color_ranges(color)
            perm = order.astype(np.int64)
            self._color_starts = starts
            self.num_colors = num_colors
            self.color_counts = color_summary(color)
        else:
            if not n_c:
                perm = np.zeros(0, np.int64)
            elif tile_mode:
                self._tile = self.tile_size or default_tile_size(n_b)
                perm = jacobi_tile_order(ra, rb, h["type"], self._tile, n_b)
            else:
                perm = jacobi_row_order(ra, rb)
            self._color_starts = np.array([0, n_c], dtype=np.int32)
            self.num_colors = 1 if n_c else 0
            self.color_counts = {}
        rows = np.ascontiguousarray(rows[perm])
        anc = np.ascontiguousarray(anc[perm])
        self._tiling = ((choose_jacobi_tiling(rows) if not tile_mode else (1, 0)) if n_c
                        else (1, 0))
        self._tile = 0
        self._tile_tabs = None
        self._rows_dev_order = rows if tile_mode else None   # for tile_pair_tables
        self._pair_dev = None
        self._cx = self._cx_dev = self._cx_struct = None
        self._plan_pf = False
        if tile_mode and n_c:
            self._tile = self.tile_size or default_tile_size(n_b)
            self._tile_tabs = tile_tables(rows, n_b, self._tile)
            ts, _, hr = self._tile_tabs
            self._cx = compact_tile_tables(rows, n_b, self._tile, ts, hr)
        self._pperm, self._pinv = pperm, pinv
        dev = self.device
        self.X = torch.from_numpy(X).to(dev)
        self.V = torch.from_numpy(V).to(dev)
        self.XP = torch.zeros((n_b, 4), dtype=torch.float32, device=dev)
        self.ACC = torch.zeros((max(n_b, 1), 4), dtype=torch.float32, device=dev)
        self.R = torch.from_numpy(rows).to(dev)
        self.ANC = torch.from_numpy(anc).to(dev)
        self._has_attach = bool(np.any(h["type"] == ATTACH)) if n_c else False
        self.LAM = torch.zeros(max(n_c, 1), dtype=torch.float32, device=dev)
        if self._tile_tabs is not None:
            self.X_B = self.X.clone()
            self.LAM_B = torch.zeros_like(self.LAM)
            ts, hs, hr = self._tile_tabs
            self._tile_dev = (torch.from_numpy(ts).to(dev), torch.from_numpy(hs).to(dev),
                              torch.from_numpy(np.ascontiguousarray(
                                  hr if len(hr) else np.zeros(1, np.int32))).to(dev))
            if self._cx is not None:
                c = self._cx
                up = lambda a, dt: torch.from_numpy(np.ascontiguousarray(  # noqa: E731
                    a if len(a) else np.zeros((1,) + a.shape[1:], dt))).to(dev)
                self._cx_dev = dict(dist_end=up(c["dist_end"], np.int32),
                                    gen_end=up(c["gen_end"], np.int32),
                                    ci=up(c["ci"], np.int32), cc=up(c["cc"], np.float32),
                                    halo2=up(c["halo2"], np.int32), pf=up(c["pf"], np.float32),
                                    pflam=torch.zeros(max(n_b, 1), dtype=torch.float32,
                                                      device=dev))
        else:
            self.X_B = self.LAM_B = None
            self._tile_dev = None
        self._perm = perm
        self.WSY = torch.empty(max(n_c, 1), dtype=torch.bfloat16, device=dev)  # warm-start y
        self.GS = None
        if self._gscale is not None:
            if not tile_mode:
                raise ValueError("a gravity scale needs solve_mode='jacobi' with the tile kernel")
            if len(self._gscale) != n_b:
                raise ValueError("gravity scale length != number of particles")
            g = self._gscale if pperm is None else self._gscale[pperm]
            self.GS = torch.from_numpy(np.ascontiguousarray(g)).to(dev)
        self.CROWS = self.CMU = None
        if self._device_contacts:
            rad = self._radius.array().astype(np.float32)
            col = self._collide.array().astype(bool)
            rad = np.where(col, rad, 0.0).astype(np.float32)
            rad_api = rad
            fr = self._friction.array().astype(np.float32)
            if pperm is not None:
                rad, fr = rad[pperm], fr[pperm]
            self._crmax = float(rad.max()) if n_b else 0.0
            # API index of each device particle, and the API pairs joined by
            # a constraint (solver.py:285-290), as sorted lo << 32 | hi keys
            self.CAPI = (torch.from_numpy(np.ascontiguousarray(pperm.astype(np.int32))).to(dev)
                         if pperm is not None else None)
            two = (h["b"] >= 0) & (h["type"] != CONTACT)
            two[two] &= (rad_api[h["a"][two]] > 0) & (rad_api[h["b"][two]] > 0)  # candidates only
            lo = np.minimum(h["a"][two], h["b"][two]).astype(np.int64)
            hi = np.maximum(h["a"][two], h["b"][two]).astype(np.int64)
            keys = np.unique((lo << 32) | hi)
            self.CEXCL = torch.from_numpy(keys).to(dev) if len(keys) else None
            self.CRAD = torch.from_numpy(np.ascontiguousarray(rad)).to(dev)
            self.CMU = torch.from_numpy(np.ascontiguousarray(fr)).to(dev)
            # the colliding particles' device indices in API order: the
            # contact search visits these only
            capi = np.nonzero(col & (rad_api > 0))[0]
            cl = (pinv[capi] if pinv is not None else capi).astype(np.int32)
            self._nclist = len(cl)
            self.CLIST = torch.from_numpy(np.ascontiguousarray(
                cl if len(cl) else np.zeros(1, np.int32))).to(dev)
            self.CWS = torch.empty(int(N.lib().xpbd_contacts_workspace_bytes(n_b)),
                                   dtype=torch.uint8, device=dev)
            self.CCOUNT = torch.zeros(1, dtype=torch.int32, device=dev)
            self._alloc_contacts(max(1024, 4 * int(col.sum())))
            if self.ACC is None or self.ACC.shape[0] < n_b:
                self.ACC = torch.zeros((max(n_b, 1), 4), dtype=torch.float32, device=dev)
            if self._tile_dev is None and self.solve_mode != "gs":   # no static rows
                self._tile = self.tile_size or default_tile_size(n_b)
                self._tile_tabs = tile_tables(np.zeros((0, 4), np.int32), n_b, self._tile)
                ts, hs, hr = self._tile_tabs
                self._tile_dev = (torch.from_numpy(ts).to(dev), torch.from_numpy(hs).to(dev),
                                  torch.zeros(1, dtype=torch.int32, device=dev))
                self.X_B = self.X.clone()
                self.LAM_B = torch.zeros_like(self.LAM)
        self.n_contacts = 0
        self._gen += 1
        self._drop_plan()
        self._dirty = False

    def _alloc_contacts(self, cap: int) -> None:
        import torch

        self._ccap = int(cap)
        dev = self.device
        if self.solve_mode == "gs":
            # Gauss-Seidel: the contact rows follow the static rows in one
            # API-ordered row set (recoloured with them every frame)
            m = len(self._rt)
            tot = m + self._ccap
            R = torch.zeros((tot, 4), dtype=torch.int32, device=dev)
            ANC = torch.zeros((tot, 4), dtype=torch.float32, device=dev)
            if m:
                R[:m] = self.R[:m]
                ANC[:m] = self.ANC[:m]
            self.R, self.ANC = R, ANC
            self.LAM = torch.zeros(max(tot, 1), dtype=torch.float32, device=dev)
            self.CROWS = self.R[m:]
            self.CLAM = self.LAM[m:tot]
            self.CINC = torch.empty(max(2 * tot, 1), dtype=torch.int32, device=dev)
            self.CCOL = torch.empty(max(tot, 1), dtype=torch.int32, device=dev)
            self.CORD = torch.empty(max(tot, 1), dtype=torch.int32, device=dev)
            self.CSTART = torch.zeros(257, dtype=torch.int32, device=dev)
            self.CNC = torch.zeros(1, dtype=torch.int32, device=dev)
            return
        self.CROWS = torch.zeros((self._ccap, 4), dtype=torch.int32, device=dev)
        self.CLAM = torch.zeros(self._ccap, dtype=torch.float32, device=dev)

    def _device_contact_pass(self) -> None:
        """Generate this frame's CONTACT rows on the device; grow the buffer
        (and drop the captured graph) if the count exceeds the capacity."""
        lib = N.lib()
        for _ in range(2):
            N.check(lib.xpbd_sphere_contacts(
                N.current_stream(), N.ptr(self.CWS), N.ptr(self.X), N.ptr(self.V),
                N.ptr(self.CRAD), N.ptr(self.CMU),
                None if self.CAPI is None else N.ptr(self.CAPI),
                N.ptr(self.CLIST), int(self._nclist),
                None if self.CEXCL is None else N.ptr(self.CEXCL),
                0 if self.CEXCL is None else int(self.CEXCL.numel()),
                self._contact_margin, self.dt, self._crmax,
                int(self.X.shape[0]), self._ccap, N.ptr(self.CROWS), N.ptr(self.CCOUNT)),
                "xpbd_sphere_contacts")
            cnt = int(self.CCOUNT.item())
            if cnt <= self._ccap:
                break
            self._alloc_contacts(int(cnt * 1.5))
            self._drop_plan()
        self.n_contacts = cnt

    def _repack_rows_inplace(self) -> None:
        h = self.host_arrays()
        ra, rb = h["a"], h["b"]
        if self._pinv is not None:
            ra = self._pinv[ra].astype(np.int32)
            rb = np.where(rb >= 0, self._pinv[np.maximum(rb, 0)], -1).astype(np.int32)
        rows, anc = pack_rows(h["type"], ra, rb, h["rest"], h["compliance"],
                              h["anchor"], h["mu"])
        if len(rows):
            m = len(rows)
            rs = np.ascontiguousarray(rows[self._perm])
            self.R[:m].copy_(self._t(rs))
            self.ANC[:m].copy_(self._t(np.ascontiguousarray(anc[self._perm])))
            if self._cx_dev is not None:   # the compact copies of the same fields
                ts, _, hr = self._tile_tabs
                c = compact_tile_tables(rs, int(self.X.shape[0]), self._tile, ts, hr)
                self._cx.update(c)
                for k in ("ci", "cc", "pf"):
                    self._cx_dev[k].copy_(self._t(np.ascontiguousarray(c[k])))

    # ---- the step (solver.py:412-467) --------------------------------------
    def _signature(self):
        return (self._gen, self.solve_mode, self.substeps, self.iterations, self.dt,
                self.gravity, self.damping, self.jacobi_relax, self.max_speed, self._tiling,
                self._tile, getattr(self, "_ccap", 0), str(self.fuse_first), self._ws_gen,
                str(self.block_frames), int(self.warm_start_every), int(self.warm_start_judge),
                str(self.tile_pair))

    # "auto" limit of the two-iteration tile launch: up to this many rows a
    # launch is latency-bound (the 128^2 cloth, 113 k rows); large scenes
    # are bandwidth-bound and the ring's extra rows would cost bytes
    PAIR_AUTO_ROWS = 400_000

    def _pair_tables(self, d, n_b, n_c):
        """Device tables of the two-iteration launch (tile_pair_tables) for
        the descriptor, when the plan can use it; None otherwise."""
        import torch

        # two plain iterations, or the learned start and one iteration
        two = self.iterations == 2 and self._ws_theta is None
        warm_one = self.iterations == 1 and self._ws_theta is not None
        if self.tile_pair is False or d.fuse_first != 2 or not (two or warm_one):
            return None
        if self.CROWS is not None or n_c == 0:
            return None
        if self.tile_pair == "auto" and n_c > self.PAIR_AUTO_ROWS:
            return None
        if self._pair_dev is None or self._pair_dev[0] != self._tile:
            t = tile_pair_tables(self._rows_dev_order, n_b, self._tile)
            dev = self.device
            up = lambda a: torch.from_numpy(np.ascontiguousarray(a if a.size else np.zeros(4, np.int32))).to(dev)  # noqa: E731
            tabs = {k: up(t[k]) for k in ("ring_start", "ring1_n", "ring_idx", "p1_start", "p1",
                                          "p2_start", "p2")}
            self._pair_dev = (self._tile, tabs, (t["lmax"], t["amax"], t["k1max"]),
                              torch.empty_like(self.XP))
        _, tabs, mx, xp_alt = self._pair_dev
        if int(N.lib().xpbd_jacobi_tile_pair_lds_bytes(*mx)) > 64 * 1024:
            return None
        return tabs, mx, xp_alt

    # "auto" limit of whole-frame workgroups in Jacobi mode: above it the
    # tile path's 256 CUs beat one CU (cloths, ms per frame, workgroup
    # against tiles: 256 particles 0.055 / 0.109, 729: 0.100 / 0.116, 1024:
    # 0.132 / 0.113)
    BLOCK_AUTO_JACOBI = 768

    def _use_block(self) -> bool:
        """Whole frames in one workgroup (``xpbd_frame_block``): small scenes
        in GS mode, or in Jacobi mode without the tile-only extensions."""
        if self.block_frames is False or self._ws_theta is not None:
            return False
        gs = self.solve_mode == "gs"
        if not gs and (self._gscale is not None or self._device_contacts):
            return False
        n = int(self.X.shape[0])
        if n > int(N.lib().xpbd_frame_block_capacity(1 if gs else 0)):
            return False
        return gs or self.block_frames is True or n <= self.BLOCK_AUTO_JACOBI

    def _drop_plan(self):
        if self._plan is not None:
            N.lib().xpbd_frame3_destroy(self._plan)
        self._plan = None
        self._plan_sig = None

    def _get_plan(self):
        sig = self._signature()
        if self._plan is not None and self._plan_sig == sig:
            return self._plan
        self._drop_plan()
        n_b = int(self.X.shape[0])
        n_c = len(self._rt)
        self._plan_pair = False
        d = N.ParticlesDesc()
        d.x, d.v, d.x_prev, d.acc = (N.ptr(self.X), N.ptr(self.V), N.ptr(self.XP),
                                     N.ptr(self.ACC))
        d.rows, d.anchors, d.lam = N.ptr(self.R), N.ptr(self.ANC), N.ptr(self.LAM)
        if self._tile_dev is not None and not self._has_attach:
            d.anchors = None        # the tile kernel variant without anchor loads
        d.n_particles, d.n_rows = n_b, n_c
        self._starts_buf = np.ascontiguousarray(self._color_starts, dtype=np.int32)
        d.color_starts = self._starts_buf.ctypes.data
        d.num_colors = int(len(self._starts_buf) - 1)
        d.mode = MODE_GS if self.solve_mode == "gs" else MODE_JACOBI
        if self._tile_dev is not None:
            d.mode = MODE_JACOBI_TILE
            d.x_alt, d.lam_alt = N.ptr(self.X_B), N.ptr(self.LAM_B)
            d.tile_row_start, d.halo_start, d.halo_rows = (N.ptr(t) for t in self._tile_dev)
            d.tile_size = self._tile
            if self.GS is not None:
                d.gscale = N.ptr(self.GS)
            if not self.fuse_first:
                d.fuse_first = 0
            elif self.fuse_first == "first" or bool(np.any(self._rt.array() == CONTACT)):
                d.fuse_first = 1    # CONTACT friction reads other tiles' x_prev
            else:
                d.fuse_first = 2
            if self.CROWS is not None:
                d.dyn_rows, d.dyn_lam = N.ptr(self.CROWS), N.ptr(self.CLAM)
                d.dyn_count, d.dyn_cap = N.ptr(self.CCOUNT), self._ccap
            pair = self._pair_tables(d, n_b, n_c)
            self._plan_pair = pair is not None
            if pair is not None:
                tabs, (lmax, amax, k1max), xp_alt = pair
                d.x_prev_alt = N.ptr(xp_alt)
                d.pair_ring_start, d.pair_ring1_n = N.ptr(tabs["ring_start"]), N.ptr(tabs["ring1_n"])
                d.pair_ring_idx = N.ptr(tabs["ring_idx"])
                d.pair_p1_start, d.pair_p1 = N.ptr(tabs["p1_start"]), N.ptr(tabs["p1"])
                d.pair_p2_start, d.pair_p2 = N.ptr(tabs["p2_start"]), N.ptr(tabs["p2"])
                d.pair_lmax, d.pair_amax, d.pair_k1max = lmax, amax, k1max
        if self._ws_theta is not None and d.mode == MODE_JACOBI_TILE:
            import torch

            self.WSI = torch.empty(int(N.lib().xpbd_mlp_image_bytes()), dtype=torch.uint8,
                                   device=self.device)
            N.check(N.lib().xpbd_mlp_pack(N.current_stream(), N.ptr(self._ws_theta),
                                          N.ptr(self.WSI)), "xpbd_mlp_pack")
            d.ws_image, d.ws_y = N.ptr(self.WSI), N.ptr(self.WSY)
            d.ws_every = int(self.warm_start_every)
            dev = self.device
            self.WSON = torch.ones(1, dtype=torch.int32, device=dev)
            self.WSST = torch.zeros(3, dtype=torch.int32, device=dev)
            d.ws_on = N.ptr(self.WSON)
            if (self.warm_start_judge and d.fuse_first >= 2 and self.substeps >= 2
                    and self.iterations >= 1 and n_c > 0):
                m = max(n_c, 1)
                f4 = lambda: torch.zeros((max(n_b, 1), 4), dtype=torch.float32, device=dev)  # noqa
                f1 = lambda k: torch.zeros(k, dtype=torch.float32, device=dev)  # noqa
                self._judge = dict(x=f4(), xa=f4(), xp=f4(), xpa=f4(), l=f1(m), la=f1(m),
                                   y=torch.zeros(m, dtype=torch.bfloat16, device=dev),
                                   rs=torch.zeros(2, dtype=torch.float64, device=dev))
                j = self._judge
                d.ws_judge = int(self.warm_start_judge)
                d.ws_jx, d.ws_jx_alt, d.ws_jxp = N.ptr(j["x"]), N.ptr(j["xa"]), N.ptr(j["xp"])
                d.ws_jxp_alt = N.ptr(j["xpa"])
                d.ws_jlam, d.ws_jlam_alt, d.ws_jy = N.ptr(j["l"]), N.ptr(j["la"]), N.ptr(j["y"])
                d.ws_rsum, d.ws_stats = N.ptr(j["rs"]), N.ptr(self.WSST)
                if self.CROWS is not None:
                    j["acc"], j["dl"] = f4(), f1(self._ccap)
                    d.ws_jacc, d.ws_jdlam = N.ptr(j["acc"]), N.ptr(j["dl"])
            else:
                self._judge = None
        self._cx_struct = None
        self._plan_pf = False
        if (d.mode == MODE_JACOBI_TILE and self._cx_dev is not None and self.compact_rows
                and n_c > 0 and self._tile >= 64):
            c = self._cx_dev
            cs = N.TileCompact()
            for k in ("dist_end", "gen_end", "ci", "cc", "halo2"):
                setattr(cs, k, N.ptr(c[k]))
            # per-particle floors unless a kernel of the plan reads the FLOOR
            # rows' multipliers from LAM (the pair launch)
            if self._cx["n_floor"] and not d.pair_p1:
                cs.pf, cs.pflam = N.ptr(c["pf"]), N.ptr(c["pflam"])
                self._plan_pf = True
            cs.general = int(self._cx["n_general"] > 0
                             or (not self._plan_pf and self._cx["n_floor"] > 0))
            if (getattr(self, "_device_contacts", False)
                    and getattr(self, "CLIST", None) is not None):
                import torch

                # only tiles holding a colliding particle read the device-row sums
                n_t = (n_b + self._tile - 1) // self._tile
                at = torch.zeros(max(n_t, 1), dtype=torch.int32, device=self.X.device)
                if self._nclist:
                    at[self.CLIST[:self._nclist].long() // self._tile] = 1
                self._acc_tiles = at
                cs.acc_tiles = N.ptr(at)
            self._cx_struct = cs
            d.compact = N.ctypes.addressof(cs)
        self._block_active = self._use_block()
        if self._block_active:
            d.compact = None
            self._plan_pf = self._plan_pair = False
            d.block = 1
            d.mode = MODE_GS if self.solve_mode == "gs" else MODE_JACOBI
            d.anchors = N.ptr(self.ANC)
            self._starts_dev = self._t(self._starts_buf)
            d.color_starts_dev = N.ptr(self._starts_dev)
            d.ws_image = d.ws_y = None
        d.substeps, d.iterations = self.substeps, self.iterations
        d.h = float(np.float32(self.dt / float(self.substeps)))
        d.gx, d.gy, d.gz = self.gravity
        d.damping, d.relax, d.max_speed = self.damping, self.jacobi_relax, self.max_speed
        d.rows_per_thread, d.window = self._tiling
        handle = N.c_void_p()
        N.check(N.lib().xpbd_frame3_create(N.ctypes.byref(d), N.ctypes.byref(handle)),
                "xpbd_frame3_create")
        self._plan = handle
        self._plan_sig = sig
        return handle

    def warm_start_stats(self) -> dict:
        """The judge of the learned start since the plan was built: judged
        frames that kept (or chose) the learned start, those that chose the
        zero start, switches between them, and whether the learned start is
        on now."""
        if self.WSST is None or self._plan is None:
            return {"judged_learned": 0, "judged_zero": 0, "switches": 0, "learned_on": None}
        st = self.WSST.cpu().numpy()
        return {"judged_learned": int(st[0]), "judged_zero": int(st[1]), "switches": int(st[2]),
                "learned_on": bool(int(self.WSON.item()))}

    def step(self, frames: int = 1) -> None:
        """Advance ``frames`` frames of ``dt`` (default one, as the reference).

        Without self-collision the constraint set is static and all frames go
        out as replays of one captured graph.  With self-collision the CONTACT
        rows are rebuilt on the host before every frame (solver.py:413-416),
        so each frame is issued eagerly."""
        per_frame = self._self_collision or self._device_contacts
        for _ in range(int(frames) if per_frame else 1):
            if self._self_collision:
                self._flush()
                self._sanitize_positions()
                self._rebuild_contacts()
            self._flush()
            if len(self._x) == 0:
                return
            if self._device_contacts:
                self._device_contact_pass()
                if self.solve_mode == "gs":
                    self._gs_device_frame()
                    continue
            plan = self._get_plan()
            n_frames = 1 if per_frame else int(frames)
            graph = int(self.use_graph and not self._self_collision)
            N.check(N.lib().xpbd_frame3_run(plan, N.current_stream(), n_frames, graph),
                    "xpbd_frame3_run")

    def _gs_device_frame(self) -> None:
        """One Gauss-Seidel frame with this frame's device CONTACT rows:
        device colouring of every row, then the frame in one workgroup."""
        lib = N.lib()
        st = N.current_stream()
        n_c = len(self._rt)
        N.check(lib.xpbd_color_rows_device(st, N.ptr(self.R), n_c, N.ptr(self.CCOUNT),
                                           self._ccap, int(self.X.shape[0]), N.ptr(self.CINC),
                                           N.ptr(self.CCOL), N.ptr(self.CORD),
                                           N.ptr(self.CSTART), N.ptr(self.CNC)),
                "xpbd_color_rows_device")
        nc = int(self.CNC.item())
        if nc < 0:
            raise RuntimeError("device colouring needed more than 256 colours")
        self.num_colors = nc
        h = float(np.float32(self.dt / float(self.substeps)))
        N.check(lib.xpbd_frame_block_ordered(
            st, N.ptr(self.X), N.ptr(self.V), N.ptr(self.XP), N.ptr(self.R), N.ptr(self.ANC),
            N.ptr(self.LAM), N.ptr(self.CSTART), N.ptr(self.CORD), N.ptr(self.CNC),
            int(self.X.shape[0]), self.substeps, self.iterations, h, *self.gravity,
            self.damping, self.jacobi_relax, self.max_speed, 1), "xpbd_frame_block_ordered")

    def device_contact_rows(self) -> dict:
        """This frame's device CONTACT rows in API particle indices: a, b,
        rest, mu (and, in Gauss-Seidel mode, their colours)."""
        cnt = min(int(self.n_contacts), getattr(self, "_ccap", 0))
        rows = self.CROWS[:cnt].cpu().numpy() if self.CROWS is not None else np.zeros((0, 4),
                                                                                      np.int32)
        a = (rows[:, 0] & 0x0FFFFFFF).astype(np.int64)
        b = rows[:, 1].astype(np.int64)
        if self._pperm is not None:
            a, b = self._pperm[a], self._pperm[b]
        out = dict(a=a, b=b, rest=rows[:, 2].copy().view(np.float32),
                   mu=rows[:, 3].copy().view(np.float32))
        if self.solve_mode == "gs" and self.CROWS is not None:
            m = len(self._rt)
            out["color"] = self.CCOL[m:m + cnt].cpu().numpy()
        return out

    def _sanitize_positions(self) -> None:
        """Reset non-finite positions to their initial scene values and zero
        their velocity (solver.py:469-483)."""
        if self.X is None:
            return
        xs = self.positions()
        if np.all(np.isfinite(xs)):
            return
        bad = np.nonzero(~np.all(np.isfinite(xs), axis=1))[0]
        safe = self._x.array().astype(np.float32)
        for i in bad.tolist():
            k = self._dev(i)
            self.X[k, :3] = self._t(safe[i])
            self.V[k, :3] = 0.0

    # ---- read-back (solver.py:487-508) ---------------------------------------
    def _api_order(self, dev_arr: np.ndarray) -> np.ndarray:
        if self._pperm is None:
            return dev_arr
        out = np.empty_like(dev_arr)
        out[self._pperm] = dev_arr
        return out

    def positions(self) -> np.ndarray:
        if self.X is None:
            return self._x.array().astype(np.float32)
        return self._api_order(self.X[:, :3].cpu().numpy())

    def velocities(self) -> np.ndarray:
        if self.V is None:
            return self._v.array().astype(np.float32)
        return self._api_order(self.V[:, :3].cpu().numpy())

    def lambdas(self) -> np.ndarray:
        """Multipliers in API row order (after the last iteration)."""
        n_c = len(self._rt)
        if self.LAM is None or self._dirty:
            return np.zeros(n_c, dtype=np.float32)
        lam = self.LAM
        launches = self.iterations + (1 if self._ws_theta is not None else 0)
        if self._tile_dev is not None and launches % 2 == 1 and not self._block_active:
            lam = self.LAM_B
        dev = lam[:n_c].cpu().numpy()
        if self._plan_pf:   # FLOOR rows kept per particle by the plan
            dev[self._cx["pf_rows"]] = self._cx_dev["pflam"].cpu().numpy()[self._cx["pf_part"]]
        out = np.zeros(n_c, dtype=np.float32)
        out[self._perm] = dev
        return out

    @property
    def num_bodies(self) -> int:
        return len(self._x)

    @property
    def num_constraints(self) -> int:
        return len(self._rt)

    def __del__(self):
        try:
            self._drop_plan()
        except Exception:
            pass


# 13 forward neighbour cell offsets (half of the 26-cell shell) plus the cell
# itself: every unordered pair of distinct or equal cells is visited once.
# The order fixes CONTACT row order and therefore coloring and GS results
# (solver.py:514-517).
_FWD_OFFSETS = (
    (1, -1, -1), (1, -1, 0), (1, -1, 1), (1, 0, -1), (1, 0, 0), (1, 0, 1),
    (1, 1, -1), (1, 1, 0), (1, 1, 1), (0, 1, -1), (0, 1, 0), (0, 1, 1), (0, 0, 1),
)


def grid_candidate_pairs(p: np.ndarray, cell: float) -> tuple[np.ndarray, np.ndarray]:
    """Candidate pairs of bodies in the same or an adjacent grid cell of side
    ``cell`` (solver.py:520-562), in the reference's enumeration order: the
    cell itself and then the 13 forward neighbour cells (``_FWD_OFFSETS``),
    each as one block; inside a block body by body in ascending index, and
    for each body the members of the target cell in ascending index (its
    own cell: only those after it).  Computed for all 14 offsets at once
    against the table of occupied cells."""
    m = p.shape[0]
    if m < 2 or cell <= 0.0:
        return np.zeros(0, np.int64), np.zeros(0, np.int64)
    ijk = np.floor(p / cell).astype(np.int64)
    ijk -= ijk.min(axis=0)
    span = int(ijk.max()) + 2              # room for the +1 neighbour on every axis
    key = (ijk[:, 0] * span + ijk[:, 1]) * span + ijk[:, 2]
    members = np.argsort(key, kind="stable")  # bodies cell by cell, ascending inside
    occupied, first, size = np.unique(key[members], return_index=True, return_counts=True)
    shift = np.array([(dx * span + dy) * span + dz for dx, dy, dz in ((0, 0, 0),) + _FWD_OFFSETS],
                     np.int64)
    target = shift[:, None] + key[None, :]     # (offset, body), offset-major
    slot = np.minimum(np.searchsorted(occupied, target), len(occupied) - 1)
    hit = occupied[slot] == target
    fan = np.where(hit, size[slot], 0).ravel()
    total = int(fan.sum())
    if total == 0:
        return np.zeros(0, np.int64), np.zeros(0, np.int64)
    owner = np.repeat(np.tile(np.arange(m, dtype=np.int64), len(shift)), fan)
    block = np.repeat(np.repeat(np.arange(len(shift)), m), fan)
    nth = np.arange(total) - np.repeat(np.cumsum(fan) - fan, fan)
    partner = members[np.repeat(first[slot].ravel(), fan) + nth]
    keep = (block > 0) | (partner > owner)
    return owner[keep], partner[keep]


def _shape_radius(shape) -> float:
    """Bounding-sphere radius used by the sphere broad phase (solver.py:565-575)."""
    if shape is None:
        return 0.0
    if shape.kind == "sphere":
        return float(shape.size[0])
    if shape.kind == "cube":
        return float(min(shape.size))
    if shape.kind == "pillar":
        return float(shape.size[0])
    return 0.0
