"""CPU-only tests of the host side: the C-ABI library's exports, the host
coloring routine (against the oracle), row packing and ordering, the host
broad phase, and the scene generators.  Nothing here launches a kernel."""
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent


def _declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        names |= set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\*?(xpbd_\w+)\s*\(", text, re.M))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    import ctypes

    from xpbd_amd import _native as N

    lib = N.lib()
    names = _declared_symbols()
    assert len(names) >= 12, names
    for name in names:
        assert hasattr(lib, name), f"{name} declared in include/ but not exported"
    assert b"gfx950" in lib.xpbd_build_info()
    assert lib.xpbd_error_string(0)
    del ctypes


def test_device_cpu_is_rejected():
    from xpbd_amd import Solver

    with pytest.raises(ValueError):
        Solver(device="cpu")


def _random_rows(seed, n_bodies=300, m=2000):
    rng = np.random.default_rng(seed)
    a = rng.integers(0, n_bodies, m)
    b = rng.integers(0, n_bodies, m)
    b = np.where(b == a, (b + 1) % n_bodies, b)
    b[rng.random(m) < 0.15] = -1
    return n_bodies, a, b


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_host_coloring_equals_oracle(oracle, seed):
    from xpbd_amd.coloring import color_constraints, color_ranges

    n, a, b = _random_rows(seed)
    got = color_constraints(n, a, b)
    want = oracle.color_constraints(n, a, b)
    assert (got == want).all()
    o1, s1, k1 = color_ranges(got)
    o2, s2, k2 = oracle.color_ranges(want)
    assert k1 == k2 and (o1 == o2).all() and (s1 == s2).all()


def test_host_coloring_dense_pile_over_64_colors(oracle):
    """A body shared by > 64 rows forces more than one 64-bit mask word."""
    from xpbd_amd.coloring import color_constraints

    m = 150
    a = np.zeros(m, np.int32)
    b = np.arange(1, m + 1, dtype=np.int32)
    got = color_constraints(m + 1, a, b)
    assert sorted(got.tolist()) == list(range(m))
    assert (got == oracle.color_constraints(m + 1, a, b)).all()


def test_host_coloring_config2_cloth(oracle):
    from xpbd_amd import Solver
    from xpbd_amd.coloring import color_constraints
    from xpbd_amd.scenes import cloth_sheet

    s, _ = cloth_sheet(Solver, res=128)
    h = s.host_arrays()
    got = color_constraints(s.num_bodies, h["a"], h["b"])
    assert (got == oracle.color_constraints(s.num_bodies, h["a"], h["b"])).all()
    assert 13 <= int(got.max()) + 1 <= 20


def test_pack_rows_layout():
    from xpbd_amd.solver import pack_rows

    rows, anc = pack_rows([0, 1, 2, 3], [5, 6, 7, 8], [9, -1, -1, 10],
                          rest=[0.5, 0.25, 0.0, 0.4], compliance=[1e-6, 2e-6, 0.0, 0.0],
                          anchor=[[0, 0, 0], [1, 2, 3], [0, 0.7, 0], [0, 0, 0]],
                          mu=[0.0, 0.0, 0.3, 0.6])
    assert (rows[:, 0] & 0x0FFFFFFF).tolist() == [5, 6, 7, 8]
    assert (rows[:, 0] >> 28).tolist() == [0, 1, 2, 3]
    assert rows[:, 1].tolist() == [9, -1, -1, 10]
    p0 = rows[:, 2].view(np.float32)
    p1 = rows[:, 3].view(np.float32)
    assert p0.tolist() == pytest.approx([0.5, 0.25, 0.7, 0.4])
    assert p1.tolist() == pytest.approx([1e-6, 2e-6, 0.3, 0.6])
    assert anc[1, :3].tolist() == [1, 2, 3]


def test_jacobi_row_order_sorted_and_tiling():
    from xpbd_amd import Solver
    from xpbd_amd.scenes import cloth_sheet
    from xpbd_amd.solver import choose_jacobi_tiling, jacobi_row_order, pack_rows

    s, _ = cloth_sheet(Solver, res=64)
    h = s.host_arrays()
    order = jacobi_row_order(h["a"], h["b"])
    key = np.where(h["b"] >= 0, np.minimum(h["a"], h["b"]), h["a"])[order]
    assert (np.diff(key) >= 0).all()
    assert sorted(order.tolist()) == list(range(len(order)))
    rows, _ = pack_rows(h["type"], h["a"], h["b"], h["rest"], h["compliance"], h["anchor"],
                        h["mu"])
    rpt, window = choose_jacobi_tiling(rows[order])
    assert rpt >= 1 and 256 <= window <= 4096


def test_config2_cloth_counts():
    from xpbd_amd import Solver
    from xpbd_amd.scenes import cloth_sheet

    s, _ = cloth_sheet(Solver, res=128)
    h = s.host_arrays()
    assert s.num_bodies == 16384
    assert int((h["type"] == 0).sum()) == 97026
    assert int((h["type"] == 2).sum()) == 16384
    assert (h["inv_mass"] == 0).sum() == 2
    c = h["compliance"][h["type"] == 0]
    assert c.min() >= 1e-8 * 0.999 and c.max() <= 1e-3 * 1.001


def _brute_pairs(p, cell):
    """All pairs whose integer cell coordinates differ by at most one per axis."""
    c = np.floor(p / cell).astype(np.int64)
    out = set()
    for i in range(len(p)):
        for j in range(i + 1, len(p)):
            if np.all(np.abs(c[i] - c[j]) <= 1):
                out.add((i, j))
    return out


@pytest.mark.parametrize("seed", [0, 1])
def test_grid_candidate_pairs_cover_neighbour_cells(seed):
    from xpbd_amd.solver import grid_candidate_pairs

    rng = np.random.default_rng(seed)
    p = rng.uniform(-1, 1, size=(150, 3)).astype(np.float32)
    a, b = grid_candidate_pairs(p, 0.3)
    got = {(min(i, j), max(i, j)) for i, j in zip(a.tolist(), b.tolist())}
    assert len(got) == len(a)                      # every unordered pair once
    assert got == _brute_pairs(p, 0.3)


def test_grid_candidate_pairs_enumeration_order():
    """Order = offset by offset ((0,0,0) first), owner ascending, partners in
    ascending index within the target cell: the CONTACT row order."""
    from xpbd_amd.solver import grid_candidate_pairs

    p = np.array([[0.05, 0.05, 0.05], [1.05, 0.05, 0.05], [0.06, 0.05, 0.05],
                  [1.06, 0.05, 0.05]], np.float32)
    a, b = grid_candidate_pairs(p, 1.0)
    # same-cell pairs first (0,2) and (1,3), then the +x neighbour pairs
    assert list(zip(a.tolist(), b.tolist())) == [(0, 2), (1, 3), (0, 1), (0, 3), (2, 1), (2, 3)]


def test_flush_preserves_reference_numerics():
    """inv_mass = float32(1/float32(m)); rest=None uses the float64 distance."""
    from xpbd_amd import Solver

    s = Solver()
    a = s.add_particle((0.0, 0.0, 0.0), mass=3.0)
    b = s.add_particle((0.1, 0.2, 0.3), mass=0.0)
    s.add_distance(a, b)
    h = s.host_arrays()
    assert h["inv_mass"][0] == np.float32(1.0) / np.float32(3.0)
    assert h["inv_mass"][1] == 0.0
    assert h["rest"][0] == np.float32(np.sqrt(0.01 + 0.04 + 0.09))


def _header_arities():
    """Parameter count of every xpbd_* function declared in include/*.h."""
    out = {}
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"(xpbd\w*)\s*\(([^;{]*?)\)\s*;", text):
            params = m.group(2).strip()
            out[m.group(1)] = 0 if params in ("", "void") else params.count(",") + 1
    return out


def test_ctypes_signatures_match_header_arity():
    from xpbd_amd import _native as N

    arity = _header_arities()
    sigs = dict(N._SIGS)
    sigs.update(N._EXTRA_SIGS)
    assert sigs, "no bound signatures"
    for name, (_, args) in sigs.items():
        assert name in arity, f"{name} bound in ctypes but not declared in include/"
        assert len(args) == arity[name], f"{name}: ctypes {len(args)} args, header {arity[name]}"


@pytest.mark.parametrize("tile,order", [(7, "lo"), (64, "lo"), (100000, "lo"), (7, "tile"),
                                        (64, "tile"), (100000, "tile")])
def test_tile_tables_cover_every_endpoint(tile, order):
    """Every (row, endpoint) pair is visited exactly once by the tile that
    owns the endpoint: as an owned row (smallest endpoint in the tile) or as
    a halo row (larger endpoint in the tile, smaller endpoint elsewhere)."""
    from xpbd_amd import Solver
    from xpbd_amd.scenes import cloth_sheet
    from xpbd_amd.solver import (jacobi_row_order, jacobi_tile_order, morton_order, pack_rows,
                                 tile_tables)

    s, _ = cloth_sheet(Solver, res=20)
    h = s.host_arrays()
    perm = morton_order(h["x"].astype(np.float64))
    inv = np.empty_like(perm)
    inv[perm] = np.arange(len(perm))
    ra = inv[h["a"]]
    rb = np.where(h["b"] >= 0, inv[np.maximum(h["b"], 0)], -1)
    rows, _ = pack_rows(h["type"], ra, rb, h["rest"], h["compliance"], h["anchor"], h["mu"])
    if order == "lo":
        rows = rows[jacobi_row_order(ra, rb)]
    else:
        perm = jacobi_tile_order(ra, rb, h["type"], tile)
        assert sorted(perm.tolist()) == list(range(len(ra)))
        rows = rows[perm]
    starts, hs, hr = tile_tables(rows, s.num_bodies, tile)
    n_t = len(starts) - 1
    seen = {}
    a = rows[:, 0] & 0x0FFFFFFF
    b = rows[:, 1]
    for t in range(n_t):
        visit = list(range(starts[t], starts[t + 1])) + hr[hs[t]:hs[t + 1]].tolist()
        for j in visit:
            for e in ([a[j]] + ([b[j]] if b[j] >= 0 else [])):
                if e // tile == t:
                    seen[(j, int(e))] = seen.get((j, int(e)), 0) + 1
    want = {(j, int(a[j])) for j in range(len(rows))} | \
           {(j, int(b[j])) for j in range(len(rows)) if b[j] >= 0}
    assert set(seen) == want and set(seen.values()) == {1}
    assert starts[0] == 0 and starts[-1] == len(rows)


def test_jacobi_tile_order_groups_by_tile_then_type_then_rank():
    """Inside each tile the rows run through the types in order, and inside a
    type through the ranks (0 = first row of a smallest endpoint), so each
    rank block holds every smallest endpoint at most once."""
    from xpbd_amd import Solver
    from xpbd_amd.scenes import cloth_sheet
    from xpbd_amd.solver import jacobi_tile_order

    s, _ = cloth_sheet(Solver, res=24)
    h = s.host_arrays()
    a, b, t = h["a"].astype(np.int64), h["b"].astype(np.int64), h["type"].astype(np.int64)
    tile = 64
    perm = jacobi_tile_order(a, b, t, tile)
    lo = np.where(b >= 0, np.minimum(a, b), a)[perm]
    tp = t[perm]
    tl = lo // tile
    assert np.all(np.diff(tl) >= 0)
    for k in np.unique(tl):
        sel = tl == k
        assert np.all(np.diff(tp[sel]) >= 0)
        for ty in np.unique(tp[sel]):
            seg = lo[sel][tp[sel] == ty]
            # ascending runs of distinct endpoints, one run per rank
            breaks = np.nonzero(np.diff(seg) <= 0)[0]
            for run in np.split(seg, breaks + 1):
                assert len(np.unique(run)) == len(run)


def test_ws_oracle_one_row_closed_form():
    """One stretched DISTANCE row between two particles with known
    velocities: C, the normal, the velocity features, lambda0 = 4 sigmoid(y)
    times the first plain step, the offsets worked out by hand, and what the
    geometry fixes: both particles move along the normal, so the new length
    is L + w lambda0, and one plain XPBD step from (x + offset, lambda0)
    solves the row exactly (C + alpha~ lambda = 0) whatever lambda0 was; a
    separated FLOOR row gets no push."""
    import struct

    from oracle import ws_ref

    h = 1.0 / 600.0
    X = np.array([[0.0, 0.0, 0.0, 2.0], [0.3, 0.4, 0.0, 1.0]])     # |d| = 0.5
    V = np.array([[0.0, 0.0, 1.0, 0.0], [0.6, 0.8, 0.0, 0.0]])
    XP = X - V * h
    comp = 1e-6
    rest = 0.4
    bits = lambda v: struct.unpack("<i", struct.pack("<f", v))[0]   # noqa: E731
    rows = np.array([[0 | (0 << 28), 1, bits(rest), bits(comp)],
                     [1 | (2 << 28), -1, bits(0.0), bits(0.5)]], np.int64).astype(np.int32)
    anc = np.zeros((2, 4))
    F = ws_ref.features(X, XP, rows, anc, h)
    rest32 = float(np.float32(rest))                           # rows hold float32
    C = 0.5 - rest32
    alpha = float(np.float32(comp)) / (h * h)
    assert F[0, 0] == 1.0 and F[1, 2] == 1.0
    assert np.isclose(F[0, 4], np.log1p(C / 1e-3))
    assert np.isclose(F[0, 6], alpha / (3.0 + alpha))
    assert np.isclose(F[1, 4], np.log1p(0.4 / 1e-3))            # floor row: C = y = 0.4
    # v_a - v_b = (-0.6, -0.8, 1): along n = (-0.6, -0.8, 0) it is 1, across 1
    assert np.isclose(F[0, 8], np.log1p(1.0 * h / 1e-3))
    assert np.isclose(F[0, 9], np.log1p(1.0 * h / 1e-3))
    assert F[0, 7] == 0.0
    step = -C / (3.0 + alpha)
    n = np.array([-0.6, -0.8, 0.0])                             # (x_a - x_b) / |d|
    for y, omega in ((0.0, 2.0), (1.0, 4.0 / (1.0 + np.exp(-1.0)))):
        Xn, lam0, acc = ws_ref.warm_launch(X, XP, rows, anc, np.array([y, 5.0]), h)
        assert np.isclose(lam0[0], omega * step)
        assert lam0[1] == 0.0                                   # separated floor row
        assert np.allclose(acc[0, :3], n * 2.0 * lam0[0]) and acc[0, 3] == 1
        assert np.allclose(acc[1, :3], -n * 1.0 * lam0[0]) and acc[1, 3] == 1
        L1 = np.linalg.norm(Xn[0, :3] - Xn[1, :3])
        assert np.isclose(L1, 0.5 + 3.0 * lam0[0])
        # one plain step from (Xn, lambda0) (kernels.py:131-160 arithmetic)
        dlam = (-(L1 - rest32) - alpha * lam0[0]) / (3.0 + alpha)
        L2 = L1 + 3.0 * dlam
        assert np.isclose((L2 - rest32) + alpha * (lam0[0] + dlam), 0.0, atol=1e-12)
        r2 = ws_ref.residual_sum(Xn, rows, anc, lam0, h)
        assert np.isclose(r2, ((L1 - rest32) + alpha * lam0[0]) ** 2)


def test_scene_files_round_trip_host(tmp_path):
    """Scene files (xpbd_amd.scene_io): a cloth and a box pile with a joint
    saved before any device work load back to identical host arrays, knobs
    and body flags; the files hold no pickled objects."""
    from xpbd_amd import Solver, Solver6DOF, scene_io
    from xpbd_amd.scenes import box_pile, cloth_sheet

    s = cloth_sheet(Solver, res=8, solve_mode="gs", device="cuda:0", substeps=7)[0]
    s.enable_self_collision(True, margin=0.01)
    p = tmp_path / "cloth.npz"
    scene_io.save(s, p)
    t = scene_io.load(p, device="cuda:0")
    h0, h1 = s.host_arrays(), t.host_arrays()
    for k in h0:
        assert np.array_equal(h0[k], h1[k]), k
    assert (t.substeps, t.solve_mode, t._self_collision) == (7, "gs", True)
    assert t._contact_margin == 0.01
    r = box_pile(Solver6DOF, 2, 2, 2, device="cuda:0", restitution=0.3)
    q = r.add_particle((0.0, 3.0, 0.0), 0.1)
    r.add_joint(r.bodies[0], q, compliance=1e-4)
    p6 = tmp_path / "pile.npz"
    scene_io.save(r, p6)
    u = scene_io.load(p6, device="cuda:0")
    h0, h1 = r.host_arrays(), u.host_arrays()
    for k in h0:
        assert np.array_equal(h0[k], h1[k]), k
    assert [b.is_particle for b in u.bodies] == [b.is_particle for b in r.bodies]
    assert (u.restitution, u.num_joints) == (0.3, 1)
    for f in (p, p6):
        with np.load(f, allow_pickle=False) as z:
            assert all(z[k].dtype != object for k in z.files)


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus above 1 exits with a message (and
    status 2) before touching a device."""
    import subprocess
    import sys
    from pathlib import Path

    root = Path(__file__).resolve().parents[1]
    r = subprocess.run([sys.executable, str(root / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


# ---- rigid host scene (xpbd_amd/solver_6dof.py) ---------------------------
def _textbook_rot(q):
    x, y, z, w = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
        [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
        [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def test_rigid_host_geometry_helpers():
    """Rotation matrices against the textbook quaternion matrix; world-AABB
    half extents |R| he; solid-box inverse inertia 12 / (m (l_j^2 + l_k^2))."""
    from xpbd_amd.solver_6dof import rotation_matrices, solid_box_inverse_inertia, world_aabb_half

    rng = np.random.default_rng(4)
    q = rng.normal(size=(20, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    R = rotation_matrices(q)
    for i in range(20):
        np.testing.assert_allclose(R[i], _textbook_rot(q[i]), atol=1e-12)
        np.testing.assert_allclose(R[i] @ R[i].T, np.eye(3), atol=1e-12)
    he = rng.uniform(0.05, 0.3, (20, 3))
    half = world_aabb_half(q, he)
    np.testing.assert_allclose(half, np.einsum("nij,nj->ni", np.abs(R), he), rtol=2e-6)
    m = rng.uniform(0.5, 3.0, 20)
    l = 2 * he
    want = np.stack([12 / (m * (l[:, 1] ** 2 + l[:, 2] ** 2)), 12 / (m * (l[:, 0] ** 2 + l[:, 2] ** 2)),
                     12 / (m * (l[:, 0] ** 2 + l[:, 1] ** 2))], 1)
    np.testing.assert_allclose(solid_box_inverse_inertia(he, m), want, rtol=1e-13)


def test_rigid_bulk_and_single_adds_build_the_same_scene():
    """add_boxes / add_joints (bulk) and add_box / add_particle / add_joint
    (one at a time) give identical host arrays, handles and flags: static
    for mass 0, no rotational inertia for point masses, rest length = the
    anchors' current separation."""
    from xpbd_amd import Solver6DOF

    rng = np.random.default_rng(1)
    n = 12
    x = rng.normal(size=(n, 3))
    he = rng.uniform(0.1, 0.3, (n, 3))
    m = rng.uniform(0.5, 2.0, n)
    m[3] = 0.0
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    a = Solver6DOF(device="cuda:0")
    for i in range(n):
        a.add_box(tuple(x[i]), tuple(he[i]), float(m[i]), quaternion=tuple(q[i]))
    p = a.add_particle((0, 2, 0), 0.1, radius=0.02)
    a.add_joint(a.bodies[0], p, compliance=1e-5, anchor_a=(0.1, 0, 0))
    a.add_joint(1, 2, rest_length=0.3)
    b = Solver6DOF(device="cuda:0")
    b.add_boxes(x, he, m, quaternions=q)
    b.add_boxes([(0, 2, 0)], 0.02, 0.1, particle=True, group=1, color=(0.85, 0.3, 0.35))
    b.add_joints([0], [n], compliance=1e-5, anchor_a=(0.1, 0, 0))
    b.add_joints([1], [2], rest_length=0.3)
    ha, hb = a.host_arrays(), b.host_arrays()
    for k in ha:
        assert np.array_equal(ha[k], hb[k]), k
    assert a.bodies == b.bodies
    assert a.bodies[3].static and ha["inv_mass"][3] == 0 and not ha["inv_I"][3].any()
    assert a.bodies[n].is_particle and not ha["inv_I"][n].any() and ha["group"][n] == 1
    rest = np.linalg.norm((x[0] + [0.1, 0, 0]) - np.array([0, 2, 0]))
    assert ha["j_rest"][0] == np.float32(rest)
    assert (a.num_bodies, a.num_joints) == (n + 1, 2)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_host_grid_broadphase_pair_set_equals_brute_force(seed):
    """The host ``grid`` broad phase (solver_6dof.py:232-258) finds exactly
    the pairs a brute-force test over all pairs keeps: per-axis overlap of
    the float32 world AABBs grown by the one global margin, not both static,
    not the same collision group > 0; each pair once."""
    from xpbd_amd.solver_6dof import host_grid_pairs, world_aabb_half

    rng = np.random.default_rng(seed)
    n = 400
    x = rng.uniform(-2, 2, (n, 3)).astype(np.float32)
    q = rng.normal(size=(n, 4))
    q = (q / np.linalg.norm(q, axis=1, keepdims=True)).astype(np.float32)
    he = rng.uniform(0.05, 0.25, (n, 3)).astype(np.float32)
    v = rng.normal(0, 1, (n, 3)).astype(np.float32)
    w = rng.normal(0, 2, (n, 3)).astype(np.float32)
    inv_m = np.where(rng.random(n) < 0.1, 0.0, 1.0).astype(np.float32)
    grp = np.where(rng.random(n) < 0.3, rng.integers(1, 3, n), 0).astype(np.int32)
    a, b, margin = host_grid_pairs(x, q, he, v, w, inv_m, grp, 1 / 60, 0.02)
    half = world_aabb_half(q, he)
    want = set()
    for i in range(n):
        for j in range(i + 1, n):
            if inv_m[i] == 0 and inv_m[j] == 0:
                continue
            if grp[i] == grp[j] and grp[i] > 0:
                continue
            if np.all(np.abs(x[i] - x[j]) - (half[i] + half[j] + np.float32(margin)) <= 0):
                want.add((i, j))
    got = [(min(i, j), max(i, j)) for i, j in zip(a.tolist(), b.tolist())]
    assert len(got) == len(set(got))
    assert set(got) == want and len(want) > 50
    vmax, wmax = float(np.abs(v).max()), float(np.abs(w).max())
    assert margin == pytest.approx(min(0.02 + 2 / 60 * (vmax + wmax * float(
        np.linalg.norm(he, axis=1).max())), float(half.max())))


@pytest.mark.parametrize("tile", [7, 64, 100000])
def test_tile_pair_tables_match_brute_force(tile):
    """The two-iteration launch's tables against set arithmetic per tile:
    ring 1 = other endpoints of the tile's rows, ring 2 = other endpoints of
    ring 1's rows outside tile and ring 1, phase-1 rows = rows incident to
    tile or ring 1 with local indices naming the right particles, phase-2
    rows = rows incident to the tile with the owner bit where the smallest
    endpoint is in the tile; LDS maxima are the per-tile maxima."""
    from xpbd_amd import Solver
    from xpbd_amd.scenes import cloth_sheet
    from xpbd_amd.solver import jacobi_tile_order, morton_order, pack_rows, tile_pair_tables

    s, _ = cloth_sheet(Solver, res=12)
    h = s.host_arrays()
    n = s.num_bodies
    perm = morton_order(h["x"].astype(np.float64))
    inv = np.empty_like(perm)
    inv[perm] = np.arange(len(perm))
    ra = inv[h["a"]]
    rb = np.where(h["b"] >= 0, inv[np.maximum(h["b"], 0)], -1)
    rows, _ = pack_rows(h["type"], ra, rb, h["rest"], h["compliance"], h["anchor"], h["mu"])
    single = rows[::5].copy()  # one-endpoint rows (floor / attach shaped)
    single[:, 1] = -1
    rows = np.concatenate([rows, single])
    a = (rows[:, 0] & 0x0FFFFFFF).astype(np.int64)
    b = rows[:, 1].astype(np.int64)
    rows = rows[jacobi_tile_order(a, b, np.zeros(len(a), np.int64), tile)]
    a = (rows[:, 0] & 0x0FFFFFFF).astype(np.int64)
    b = rows[:, 1].astype(np.int64)
    t = tile_pair_tables(rows, n, tile)
    ends = [{int(a[j])} | ({int(b[j])} if b[j] >= 0 else set()) for j in range(len(rows))]

    def incident(ps):
        return {j for j in range(len(rows)) if ends[j] & ps}

    n_t = (n + tile - 1) // tile
    lmax = amax = k1max = 0
    for k in range(n_t):
        T = set(range(k * tile, min((k + 1) * tile, n)))
        rows_t = incident(T)
        r1 = set().union(*(ends[j] for j in rows_t)) - T if rows_t else set()
        rows_r1 = incident(r1)
        r2 = (set().union(*(ends[j] for j in rows_r1)) if rows_r1 else set()) - T - r1
        rb0, re0 = t["ring_start"][k], t["ring_start"][k + 1]
        ring = t["ring_idx"][rb0:re0].tolist()
        n1 = int(t["ring1_n"][k])
        assert set(ring[:n1]) == r1 and set(ring[n1:]) == r2 and len(ring) == len(r1) + len(r2)

        def glob(l):
            return k * tile + l if l < tile else ring[l - tile]

        p1 = t["p1"][t["p1_start"][k]:t["p1_start"][k + 1]]
        assert sorted(p1[:, 0].tolist()) == sorted(rows_t | rows_r1)
        for j, la, lb, _ in p1.tolist():
            assert glob(la) == a[j] and (lb == -1 if b[j] < 0 else glob(lb) == b[j])
        p2 = t["p2"][t["p2_start"][k]:t["p2_start"][k + 1]].astype(np.int64)
        pos = p2 & 0x7FFFFFFF
        own = p2 < 0
        assert sorted(p1[pos, 0].tolist()) == sorted(rows_t)
        lo = np.where(b >= 0, np.minimum(a, b), a)
        assert (own == (lo[p1[pos, 0]] // tile == k)).all()
        lmax = max(lmax, tile + len(r1) + len(r2))
        amax = max(amax, tile + len(r1))
        k1max = max(k1max, len(rows_t | rows_r1))
    amax += amax & 1          # rounded up to even (16-byte LDS records follow the sums)
    assert (t["lmax"], t["amax"], t["k1max"]) == (max(lmax, amax), amax, k1max)


def _tile_rows(s, tile):
    """Rows of a host scene in the fused kernel's order, with its tables."""
    from xpbd_amd.solver import jacobi_tile_order, morton_order, pack_rows, tile_tables

    h = s.host_arrays()
    perm = morton_order(h["x"].astype(np.float64))
    inv = np.empty_like(perm)
    inv[perm] = np.arange(len(perm))
    ra = inv[h["a"]]
    rb = np.where(h["b"] >= 0, inv[np.maximum(h["b"], 0)], -1)
    rows, _ = pack_rows(h["type"], ra, rb, h["rest"], h["compliance"], h["anchor"], h["mu"])
    rows = rows[jacobi_tile_order(ra, rb, h["type"], tile, s.num_bodies)]
    return rows, tile_tables(rows, s.num_bodies, tile)


@pytest.mark.parametrize("tile", [32, 64, 1024])
def test_compact_tile_tables_describe_the_same_rows(tile):
    """Each tile's rows split into compact DISTANCE rows (their (lo, hi)
    decode back to the row's endpoints, rest and compliance are the row's),
    rows read whole (ATTACH, and the FLOOR rows of a particle with two) and
    single FLOOR rows, whose floor and friction land in the per-particle
    table; halo entries name the owner tile's base for compact rows only."""
    from xpbd_amd import Body, Solver
    from xpbd_amd.scenes import cloth_sheet
    from xpbd_amd.solver import compact_tile_tables

    s, _ = cloth_sheet(Solver, res=20)
    for i in range(0, 400, 37):
        s.add_attach(Body(i), (0.1 * i, 1.0, 0.0), compliance=1e-4, rest=0.05)
    s.add_floor_contact(Body(5), floor_y=-0.2, friction=0.3)      # particle 5: two FLOOR rows
    rows, (ts, hs, hr) = _tile_rows(s, tile)
    n = s.num_bodies
    c = compact_tile_tables(rows, n, tile, ts, hr)
    t = rows[:, 0] >> 28
    a = (rows[:, 0] & 0x0FFFFFFF).astype(np.int64)
    b = rows[:, 1].astype(np.int64)
    lo = np.where(b >= 0, np.minimum(a, b), a)
    hi = np.where(b >= 0, np.maximum(a, b), a)
    nfl = np.bincount(a[t == 2], minlength=n)
    seen_floor = []
    for k in range(len(ts) - 1):
        r0, r1, r2, r3 = ts[k], c["dist_end"][k], c["gen_end"][k], ts[k + 1]
        assert r0 <= r1 <= r2 <= r3
        base = k * tile
        for j in range(r0, r1):
            q = int(c["ci"][j, 0]) & 0xFFFFFFFF
            assert t[j] == 0
            assert (base + (q & 1023), base + (q >> 10)) == (lo[j], hi[j])
            assert c["ci"][j, 1] == rows[j, 2]
            assert c["cc"][j].view(np.int32) == rows[j, 3]
        for j in range(r1, r2):
            assert t[j] != 0 and not (t[j] == 2 and nfl[a[j]] == 1)
        for j in range(r2, r3):
            assert t[j] == 2 and nfl[a[j]] == 1
            assert c["pf"][a[j], 0].view(np.int32) == rows[j, 2]
            assert c["pf"][a[j], 1].view(np.int32) == rows[j, 3]
            seen_floor.append(a[j])
    assert sorted(seen_floor) == sorted(np.nonzero(nfl == 1)[0].tolist())
    assert np.isneginf(c["pf"][nfl != 1, 0]).all()
    assert (c["pf_part"] == a[c["pf_rows"]]).all()
    assert (c["halo2"][:, 0] == hr).all()
    own = (lo[hr] // tile) * tile
    assert (c["halo2"][:, 1] == np.where(t[hr] == 0, own, -1)).all()
    assert np.count_nonzero(t == 1) > 0 and nfl[nfl > 0].max() == 2


def test_compact_rows_need_a_22_bit_span():
    """A DISTANCE row whose larger endpoint lies 2^22 or more particles past
    its tile's base is read whole; one just inside the span is compact."""
    from xpbd_amd.solver import CLS_COMPACT, CLS_GENERAL, tile_row_class

    tile = 512
    a = np.array([0, 1, 700, 0])
    b = np.array([(1 << 22) - 1, 1 << 22, 512 + (1 << 22) - 1, -1])
    t = np.array([0, 0, 0, 1])
    cls = tile_row_class(a, b, t, tile, 5_000_000)
    assert cls.tolist() == [CLS_COMPACT, CLS_GENERAL, CLS_COMPACT, CLS_GENERAL]
    assert (tile_row_class(a, b, t, 2048, 5_000_000) == CLS_GENERAL).all()
