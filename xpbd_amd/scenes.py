"""Synthetic scene generators for the benchmark configurations and tests.

Each builder drives the public API (``add_particle``/``add_distance``/...)
the way the reference's example scripts do (examples/hanging_chain.py:18-30,
examples/cloth_drop.py:18-46, examples/benchmark.py:71-107), with the sizes of
the workloads below:

* ``hanging_chain``  — 64 particles of 1 kg spaced 0.05 m along +x, particle 0
  static, 63 DISTANCE rows (compliance 1e-6), jitter N(0, 1e-4) m, GS.
* ``cloth_sheet``    — res x res grid at 1 cm spacing, 1e-3 kg particles, two
  pinned corners, structural + both shear diagonals + bend DISTANCE rows
  (compliance log-uniform in [1e-8, 1e-3]), a FLOOR row per particle at y=0
  with mu=0.4, jitter N(0, 1e-4) m.  res=128 gives 97 026 DISTANCE rows +
  16 384 FLOOR rows.
* ``cloth_bench``    — the reference benchmark cloth (structural + one shear
  diagonal, examples/benchmark.py:71-90).
"""
from __future__ import annotations

import numpy as np

from .scene import Body


def hanging_chain(solver_cls, n: int = 64, spacing: float = 0.05, compliance: float = 1e-6,
                  seed: int = 0, jitter: float = 1e-4, **kw):
    kw.setdefault("dt", 1.0 / 60.0)
    kw.setdefault("substeps", 15)
    kw.setdefault("iterations", 2)
    kw.setdefault("solve_mode", "gs")
    s = solver_cls(**kw)
    rng = np.random.default_rng(seed)
    jit = rng.normal(0.0, jitter, size=(n, 3))
    prev = None
    for i in range(n):
        p = (i * spacing + jit[i, 0], jit[i, 1], jit[i, 2])
        b = s.add_particle(p, mass=0.0 if i == 0 else 1.0, collide=False)
        if prev is not None:
            s.add_distance(prev, b, rest=spacing, compliance=compliance)
        prev = b
    return s


def cloth_edges(res: int) -> tuple[np.ndarray, np.ndarray]:
    """Edge list of the config-2 cloth in emission order: per grid cell
    (iy, ix) row-major, structural-x, structural-y, shear, shear, bend-x,
    bend-y (the edges examples/cloth_drop.py:35-45 builds)."""
    k = np.arange(res * res, dtype=np.int64)
    ix = k % res
    iy = k // res
    ea = np.empty((res * res, 6), np.int64)
    eb = np.empty((res * res, 6), np.int64)
    valid = np.empty((res * res, 6), bool)
    ea[:, 0], eb[:, 0], valid[:, 0] = k, k + 1, ix + 1 < res
    ea[:, 1], eb[:, 1], valid[:, 1] = k, k + res, iy + 1 < res
    ea[:, 2], eb[:, 2], valid[:, 2] = k, k + res + 1, (ix + 1 < res) & (iy + 1 < res)
    ea[:, 3], eb[:, 3], valid[:, 3] = k + 1, k + res, (ix + 1 < res) & (iy + 1 < res)
    ea[:, 4], eb[:, 4], valid[:, 4] = k, k + 2, ix + 2 < res
    ea[:, 5], eb[:, 5], valid[:, 5] = k, k + 2 * res, iy + 2 < res
    return ea[valid], eb[valid]


def cloth_sheet(solver_cls, res: int = 128, spacing: float = 0.01, mass: float = 1e-3,
                height: float = 0.5, seed: int = 0, jitter: float = 1e-4,
                compliance_range=(1e-8, 1e-3), floor_mu: float = 0.4, with_floor: bool = True,
                n_sheets: int = 1, **kw):
    """Cloth of config 2 (``n_sheets`` independent copies side by side, for
    the HBM-roofline scene): particles row-major per sheet, DISTANCE rows per
    ``cloth_edges`` with rest = current separation, then one FLOOR row per
    particle."""
    kw.setdefault("dt", 1.0 / 60.0)
    kw.setdefault("substeps", 15)
    kw.setdefault("iterations", 2)
    s = solver_cls(**kw)
    rng = np.random.default_rng(seed)
    iy, ix = np.meshgrid(np.arange(res), np.arange(res), indexing="ij")
    base_pos = np.stack([(ix.ravel() - (res - 1) / 2) * spacing,
                         np.full(res * res, float(height)),
                         (iy.ravel() - (res - 1) / 2) * spacing], axis=1)
    pinned = np.zeros(res * res, bool)
    pinned[0] = pinned[res - 1] = True
    masses = np.where(pinned, 0.0, mass)
    ea, eb = cloth_edges(res)
    lo, hi = np.log(compliance_range[0]), np.log(compliance_range[1])
    for sheet in range(n_sheets):
        pos = base_pos + rng.normal(0.0, jitter, size=base_pos.shape)
        pos[:, 0] += sheet * (res + 4) * spacing
        first = s.add_particles(pos, masses)[0]
    grid = np.arange(res * res, dtype=np.int64).reshape(res, res)
    for sheet in range(n_sheets):
        off = sheet * res * res
        comp = np.exp(rng.uniform(lo, hi, size=ea.shape[0]))
        s.add_distances(ea + off, eb + off, rest=None, compliance=comp)
    if with_floor:
        s.add_floor_contacts(np.arange(n_sheets * res * res), floor_y=0.0, friction=floor_mu)
    del first
    return s, grid


def cloth_bench(solver_cls, res: int, mode: str, substeps: int = 15, iterations: int = 2, **kw):
    """examples/benchmark.py:71-90 cloth: structural edges + one shear diagonal."""
    s = solver_cls(dt=1 / 60, substeps=substeps, iterations=iterations, solve_mode=mode,
                   damping=0.01, **kw)
    h = 2.4 / (res - 1)
    grid = np.empty((res, res), np.int64)
    for iy in range(res):
        for ix in range(res):
            pinned = iy == 0 and ix in (0, res - 1)
            b = s.add_particle(((ix - res / 2) * h, 2.5, (iy - res / 2) * h),
                               mass=0.0 if pinned else 0.02, collide=False)
            grid[iy, ix] = b.index
    for iy in range(res):
        for ix in range(res):
            if ix + 1 < res:
                s.add_distance(Body(int(grid[iy, ix])), Body(int(grid[iy, ix + 1])), compliance=1e-6)
            if iy + 1 < res:
                s.add_distance(Body(int(grid[iy, ix])), Body(int(grid[iy + 1, ix])), compliance=1e-6)
            if ix + 1 < res and iy + 1 < res:
                s.add_distance(Body(int(grid[iy, ix])), Body(int(grid[iy + 1, ix + 1])),
                               compliance=2e-6)
    return s


def sphere_stack(solver_cls, n_side: int, rows: int, mode: str, substeps: int = 15,
                 iterations: int = 2, **kw):
    """examples/benchmark.py:92-107: rows of n_side^2 spheres (r = 0.16 m,
    1 kg, mu 0.4) on a floor, self-collision on."""
    from .scene import Shape

    s = solver_cls(dt=1 / 60, substeps=substeps, iterations=iterations, solve_mode=mode, **kw)
    rng = np.random.default_rng(0)
    r = 0.16
    for ly in range(rows):
        for ix in range(n_side):
            for iz in range(n_side):
                b = s.add_particle(((ix - n_side / 2) * 2.3 * r + rng.uniform(-2e-3, 2e-3),
                                    r + ly * 2.3 * r,
                                    (iz - n_side / 2) * 2.3 * r + rng.uniform(-2e-3, 2e-3)),
                                   mass=1.0, shape=Shape("sphere", (r,)), friction=0.4)
                s.add_floor_contact(b, floor_y=r, friction=0.4)
    s.enable_self_collision(True, default_friction=0.4)
    return s


def _part1by1(v):
    """Spread the low 16 bits of an int64 tensor to even bit positions."""
    v = v & 0xFFFF
    v = (v | (v << 8)) & 0x00FF00FF
    v = (v | (v << 4)) & 0x0F0F0F0F
    v = (v | (v << 2)) & 0x33333333
    v = (v | (v << 1)) & 0x55555555
    return v


def packed_cloth_device(res: int, n_sheets: int, device, seed: int = 0, spacing: float = 0.01,
                        mass: float = 1e-3, height: float = 0.5, floor_mu: float = 0.4,
                        compliance_range=(1e-8, 1e-3), tile: int = 512):
    """The config-2 cloth generator scaled to ``n_sheets`` sheets of res^2,
    built directly on the device in the packed layout of include/xpbd.h, for
    the HBM-roofline scene (8 x 1024^2 = 58.6M rows), where host-side scene
    lists would cost minutes.  The edge set per particle k (grid (ix, iy)) is
    that of ``cloth_edges``: (k,+x), (k,+y), (k,+x+y), (k+x,+y-x) as
    (k,k+res-1) from its right neighbour, (k,+2x), (k,+2y), plus FLOOR(k).

    Particles are numbered in 2-D Z-order within each sheet (the locality
    layout the solver's Jacobi path uses), rows are in the fused kernel's
    order (``solver.jacobi_tile_order``), and its tile tables for ``tile``
    particles per tile are built on the device.

    Returns dict of torch tensors X, X_B, V, XP, ACC (float4 per particle), R
    (int4 rows), ANC, LAM, LAM_B, TRS, HS, HR (tile tables), CX (the compact
    tables, include/xpbd.h xpbd_tile_compact) and counts.
    ``height`` <= 0 lays the sheets on the floor (every FLOOR row active,
    x_prev = x so friction reads it)."""
    import torch

    g = torch.Generator(device=device)
    g.manual_seed(seed)
    n1 = res * res
    n = n1 * n_sheets
    k = torch.arange(n, device=device, dtype=torch.int64)
    kl = k % n1
    ix = kl % res
    iy = kl // res
    sheet = k // n1
    # Z-order numbering: new index of row-major particle k
    code = sheet * (1 << 40) + (_part1by1(iy) << 1 | _part1by1(ix))
    order = torch.argsort(code, stable=True)          # order[new] = old
    inv = torch.empty_like(order)
    inv[order] = k
    Xo = torch.empty((n, 4), device=device, dtype=torch.float32)
    Xo[:, 0] = (ix.float() - (res - 1) / 2) * spacing + sheet.float() * (res + 4) * spacing
    Xo[:, 1] = height
    Xo[:, 2] = (iy.float() - (res - 1) / 2) * spacing
    Xo[:, :3] += torch.randn((n, 3), device=device, generator=g) * 1e-4
    pinned = (iy == 0) & ((ix == 0) | (ix == res - 1))
    w = float(torch.tensor(1.0 / mass, dtype=torch.float32))
    Xo[:, 3] = torch.where(pinned, torch.zeros_like(Xo[:, 3]), torch.full_like(Xo[:, 3], w))
    nb = torch.stack([k + 1, k + res, k + res + 1, k + res - 1, k + 2, k + 2 * res], dim=1)
    valid = torch.stack([ix + 1 < res, iy + 1 < res, (ix + 1 < res) & (iy + 1 < res),
                         (ix >= 1) & (iy + 1 < res), ix + 2 < res, iy + 2 < res], dim=1)
    kinds = torch.cat([valid, torch.ones((n, 1), dtype=torch.bool, device=device)], dim=1)
    bcol = torch.cat([nb, torch.full((n, 1), -1, device=device, dtype=torch.int64)], dim=1)
    a = k[:, None].expand(n, 7)[kinds]
    b = bcol[kinds]
    types = torch.zeros((n, 7), dtype=torch.int64, device=device)
    types[:, 6] = 2
    t = types[kinds]
    del nb, valid, kinds, bcol, types
    m = a.shape[0]
    is_d = t == 0
    bb = torch.where(is_d, b, a)
    rest = torch.linalg.vector_norm(Xo[a, :3] - Xo[bb, :3], dim=1)
    lo_c, hi_c = torch.log(torch.tensor(compliance_range, dtype=torch.float64))
    comp = torch.exp(torch.rand(m, device=device, generator=g, dtype=torch.float64)
                     * (hi_c - lo_c) + lo_c).float()
    # renumber endpoints; order rows as the fused tile kernel wants them:
    # by tile of the smallest endpoint, then type, then rank among the rows of
    # the same smallest endpoint, then that endpoint
    a = inv[a]
    b = torch.where(is_d, inv[torch.clamp(b, min=0)], b)
    lo = torch.where(is_d, torch.minimum(a, b), a)
    hi = torch.where(is_d, torch.maximum(a, b), a)
    ro = torch.argsort(lo, stable=True)
    a, b, t, lo, hi, rest, comp, is_d = (v[ro] for v in (a, b, t, lo, hi, rest, comp, is_d))
    rank = torch.arange(m, device=device) - torch.searchsorted(lo, lo)
    key = (lo // tile) * (1 << 32) + t * (1 << 28) + rank * (1 << 16) + lo % tile
    ro = torch.argsort(key, stable=True)
    del rank, key
    a, b, t, lo, hi, rest, comp, is_d = (v[ro] for v in (a, b, t, lo, hi, rest, comp, is_d))
    p0 = torch.where(is_d, rest, torch.zeros_like(rest))
    p1 = torch.where(is_d, comp, torch.full_like(comp, floor_mu))
    R = torch.empty((m, 4), device=device, dtype=torch.int32)
    R[:, 0] = (a | (t << 28)).to(torch.int32)
    R[:, 1] = b.to(torch.int32)
    R[:, 2] = p0.view(torch.int32)
    R[:, 3] = p1.view(torch.int32)
    n_tiles = (n + tile - 1) // tile
    TRS = torch.searchsorted(lo // tile, torch.arange(n_tiles + 1, device=device)).to(torch.int32)
    TRS[-1] = m
    cross = torch.nonzero((hi // tile) != (lo // tile)).flatten()
    ht = hi[cross] // tile
    so = torch.argsort(ht, stable=True)
    HR = cross[so].to(torch.int32)
    HS = torch.searchsorted(ht[so].contiguous(), torch.arange(n_tiles + 1, device=device)
                            ).to(torch.int32)
    n_dist = int(is_d.sum())
    X = Xo[order].contiguous()
    # compact tables (xpbd_tile_compact): every DISTANCE row is compact (a
    # sheet spans 2^20 particles), every particle has one FLOOR row, so a
    # tile's rows are its DISTANCE rows then its FLOOR rows
    base = (lo // tile) * tile
    assert int((hi - base).max()) < (1 << 22)
    CI = torch.stack([torch.where(is_d, (lo - base) | ((hi - base) << 10),
                                  torch.zeros_like(lo)).to(torch.int32), R[:, 2]], 1).contiguous()
    CC = R[:, 3].view(torch.float32).clone()
    DE = (TRS[:-1].long() + torch.bincount((lo // tile)[is_d], minlength=n_tiles)).to(torch.int32)
    H2 = torch.stack([cross[so].to(torch.int32), base[cross[so]].to(torch.int32)], 1).contiguous()
    PF = torch.empty((n, 2), device=device, dtype=torch.float32)
    PF[a[~is_d], 0] = p0[~is_d]
    PF[a[~is_d], 1] = p1[~is_d]
    del a, b, t, lo, hi, rest, comp, p0, p1, is_d, Xo, ro, cross, ht, so, base
    return dict(
        X=X, X_B=X.clone(), V=torch.zeros((n, 4), device=device),
        XP=X.clone() if height <= 0 else torch.zeros((n, 4), device=device),
        ACC=torch.zeros((n, 4), device=device), R=R,
        # no ATTACH rows: the anchor array is never read, one slot suffices
        ANC=torch.zeros((1, 4), device=device), LAM=torch.zeros(m, device=device),
        LAM_B=torch.zeros(m, device=device), TRS=TRS, HS=HS,
        HR=HR if HR.numel() else torch.zeros(1, dtype=torch.int32, device=device),
        CX=dict(dist_end=DE, gen_end=DE.clone(), ci=CI, cc=CC,
                halo2=H2 if H2.numel() else torch.zeros((1, 2), dtype=torch.int32, device=device),
                pf=PF, pflam=torch.zeros(n, device=device), general=0),
        n_particles=n, n_rows=m, n_distance=n_dist, tile=tile, n_halo=int(HR.numel()))


def _random_quats(rng, n):
    """Uniform random unit quaternions (Shoemake 1992), xyzw."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    a, b = np.sqrt(1 - u1), np.sqrt(u1)
    return np.stack([a * np.sin(2 * np.pi * u2), a * np.cos(2 * np.pi * u2),
                     b * np.sin(2 * np.pi * u3), b * np.cos(2 * np.pi * u3)], axis=1)


def box_pile(solver_cls, nx, ny, nz, seed=0, **kw):
    """examples/benchmark.py:23-34: an nx x ny x nz pile of 0.16 m cubes of 1 kg."""
    kw.setdefault("dt", 1 / 60)
    kw.setdefault("friction", 0.6)
    s = solver_cls(**kw)
    h = 0.16
    rng = np.random.default_rng(seed)
    sp = 2.25 * h
    for ly in range(ny):
        for ix in range(nx):
            for iz in range(nz):
                s.add_box((ix * sp + rng.uniform(-1e-3, 1e-3), h + ly * sp,
                           iz * sp + rng.uniform(-1e-3, 1e-3)), (h, h, h), mass=1.0)
    return s


def unified_scene(solver_cls, seed=0, lattice=(12, 10, 12), cloth_res=32, chain_links=20,
                  **kw):
    """Config 3: 1440 boxes with half extents U(0.1, 0.3) m and density
    1000 kg/m^3 on a jittered 12 x 12 x 10 lattice above the floor with
    uniform random orientations (mu 0.5, restitution 0.2), plus a 32 x 32
    particle cloth hung from its top row by compliant joints and a 20-link
    joint chain (compliance 1e-5) hung from a static anchor."""
    kw.setdefault("dt", 1 / 60)
    kw.setdefault("substeps", 15)
    kw.setdefault("iterations", 2)
    kw.setdefault("friction", 0.5)
    kw.setdefault("restitution", 0.2)
    s = solver_cls(**kw)
    rng = np.random.default_rng(seed)
    nx, ny, nz = lattice
    sp = 1.1
    quats = _random_quats(rng, nx * ny * nz)
    k = 0
    for iy in range(ny):
        for ix in range(nx):
            for iz in range(nz):
                he = rng.uniform(0.1, 0.3, 3)
                m = 8.0 * 1000.0 * he[0] * he[1] * he[2]
                p = (ix * sp + rng.uniform(-0.05, 0.05), 0.6 + iy * sp,
                     iz * sp + rng.uniform(-0.05, 0.05))
                s.add_box(p, tuple(he), mass=float(m), quaternion=tuple(quats[k]))
                k += 1
    # cloth curtain beside the pile
    R = cloth_res
    cs = 1.6 / (R - 1)
    idx = np.empty((R, R), np.int64)
    for iy in range(R):
        for ix in range(R):
            pinned = iy == 0
            b = s.add_particle((-3.0, 4.0 - iy * cs, ix * cs), mass=0.0 if pinned else 0.02,
                               radius=0.012, group=1, static=pinned)
            idx[iy, ix] = b.index
    for iy in range(R):
        for ix in range(R):
            if ix + 1 < R:
                s.add_joint(int(idx[iy, ix]), int(idx[iy, ix + 1]), compliance=1e-7)
            if iy + 1 < R:
                s.add_joint(int(idx[iy, ix]), int(idx[iy + 1, ix]), compliance=1e-7)
            if ix + 1 < R and iy + 1 < R:
                s.add_joint(int(idx[iy, ix]), int(idx[iy + 1, ix + 1]), compliance=2e-7)
    # compliant link chain from a static anchor
    hl = 0.08
    prev = s.add_box((-6.0, 5.0, 0.0), (0.05, 0.05, 0.05), mass=0.0, static=True)
    for i in range(chain_links):
        b = s.add_box((-6.0 + (i + 1) * 2 * hl, 5.0, 0.0), (hl, 0.03, 0.03), mass=0.5)
        anchor_prev = (0.0, 0.0, 0.0) if i == 0 else (hl, 0.0, 0.0)
        s.add_joint(prev, b, compliance=1e-5, rest_length=0.0, anchor_a=anchor_prev,
                    anchor_b=(-hl, 0.0, 0.0))
        prev = b
    return s
