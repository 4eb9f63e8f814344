"""GPU parity tests of the 3-DOF solver: the HIP path (through the C-ABI)
against the CPU oracle on identical seeded inputs, plus the reference's own
physics-invariant tests (tests/test_solver.py) run on the device.

Tolerances (array-level relative error max|gpu - oracle| / max|oracle|):
  GS      1e-5 per frame (the kernels follow the reference's FP32 operation
          order and colors touch disjoint particles, so GS is deterministic)
  Jacobi  1e-4 per substep from identical inputs (atomic accumulation order
          differs), a bounded 8-frame trajectory drift, and RMS constraint
          violation / kinetic energy within 1 % over 600 frames.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
G = (0.0, -9.81, 0.0)


def _rel(g, o):
    g = np.asarray(g, np.float64)
    o = np.asarray(o, np.float64)
    den = max(float(np.abs(o).max()), 1e-30)
    return float(np.abs(g - o).max()) / den


def _oracle_scene(oracle, s):
    return oracle.Scene3(s.host_arrays(), dt=s.dt, substeps=s.substeps, iterations=s.iterations,
                         gravity=s.gravity, mode=s.solve_mode, damping=s.damping,
                         relax=s.jacobi_relax, max_speed=s.max_speed)


def _compare_frames(oracle, s, frames, tol, every=1, lam=True):
    orc = _oracle_scene(oracle, s)
    worst = 0.0
    exact = True
    for f in range(frames):
        s.step()
        orc.step(1)
        if (f + 1) % every and f + 1 != frames:
            continue
        xg, vg = s.positions(), s.velocities()
        errs = [_rel(xg, orc.x), _rel(vg, orc.v)]
        exact &= bool((xg == orc.x).all() and (vg == orc.v).all())
        if lam and s.num_constraints:
            lg = s.lambdas()
            errs.append(_rel(lg, orc.lam))
            exact &= bool((lg == orc.lam).all())
        e = max(errs)
        worst = max(worst, e)
        assert e <= tol, f"frame {f}: rel err {e:.3e} > {tol}"
    return worst, exact


def test_chain_gs_matches_oracle_600_frames(oracle):
    """Config 1: 64-particle hanging chain, GS, 15x2, 600 frames."""
    from xpbd_amd import Solver
    from xpbd_amd.scenes import hanging_chain

    s = hanging_chain(Solver, device=DEV)
    worst, exact = _compare_frames(oracle, s, 600, 1e-5)
    print(f"chain gs: worst rel err {worst:.3e}, bit-exact={exact}")
    assert s.num_colors == 2


def _floor_rows_active(s, lam):
    """Number of FLOOR rows (API order) with a nonzero multiplier."""
    return int((lam[s.host_arrays()["type"] == 2] != 0).sum())


def test_cloth_config2_gs_matches_oracle(oracle):
    """Config 2: 128^2 cloth, 97 026 DISTANCE + 16 384 FLOOR rows, GS 15x2,
    compared frame by frame over 50 frames (750 substeps): the sheet swings
    down from y = 0.5 m and lands on the floor around frame 20, so the last
    frames run with FLOOR rows and their friction active."""
    from xpbd_amd import Solver
    from xpbd_amd.scenes import cloth_sheet

    s, _ = cloth_sheet(Solver, res=128, solve_mode="gs", device=DEV)
    worst, exact = _compare_frames(oracle, s, 50, 1e-5, every=5)
    active = _floor_rows_active(s, s.lambdas())
    print(f"cloth128 gs: worst rel err {worst:.3e}, bit-exact={exact}, "
          f"{active} FLOOR rows active at frame 50")
    assert active > 100


def _substep_check(oracle, build, n_checks, frames_between, tol, start=0, floor_active=0):
    """Per-substep parity from identical inputs: the oracle runs the scene
    forward (``start`` frames, then ``frames_between`` per check); at each
    check point its state is loaded into a one-substep GPU solver and both
    advance one substep (same h, same rows).  ``floor_active``: at least
    that many FLOOR rows must carry a nonzero multiplier at every check."""
    s = build(1)
    s_full = build(None)
    orc = _oracle_scene(oracle, s_full)
    one = _oracle_scene(oracle, s)
    s._flush()
    worst = 0.0
    orc.step(start)
    for k in range(n_checks):
        orc.step(frames_between)
        one.x[:] = orc.x
        one.v[:] = orc.v
        s.set_state(orc.x, orc.v)
        s.step()
        one.step(1)
        errs = [_rel(s.positions(), one.x), _rel(s.velocities(), one.v),
                _rel(s.lambdas(), one.lam)]
        worst = max(worst, max(errs))
        assert max(errs) <= tol, f"check {k}: x/v/lam rel err {errs}"
        if floor_active:
            assert _floor_rows_active(s, one.lam) >= floor_active, k
    return worst


@pytest.mark.parametrize("start", [0, 40])
def test_cloth_config2_jacobi_per_substep(oracle, start):
    """Config 2 in Jacobi mode: per substep from identical inputs, positions,
    velocities and lambda agree to 1e-4 (atomic order differs); from frame
    40 on the sheet lies on the floor, FLOOR rows and friction active."""
    from xpbd_amd import Solver
    from xpbd_amd.scenes import cloth_sheet

    def build(sub):
        kw = dict(res=128, solve_mode="jacobi", device=DEV)
        if sub == 1:
            h = float(np.float32(1.0 / 60.0 / 15.0))
            kw.update(dt=h, substeps=1)
        return cloth_sheet(Solver, **kw)[0]

    worst = _substep_check(oracle, build, n_checks=5, frames_between=3, tol=1e-4, start=start,
                           floor_active=100 if start else 0)
    print(f"cloth128 jacobi per-substep from frame {start}: worst rel err {worst:.3e}")


def test_cloth_config2_jacobi_trajectory(oracle):
    """Config 2 Jacobi over 8 frames (120 substeps) without re-synchronising:
    the atomic-order rounding differences stay bounded (velocity is a finite
    difference over h, so one ulp of x is ~1e-4 of v)."""
    from xpbd_amd import Solver
    from xpbd_amd.scenes import cloth_sheet

    s, _ = cloth_sheet(Solver, res=128, solve_mode="jacobi", device=DEV)
    worst, _ = _compare_frames(oracle, s, 8, 2e-3)
    print(f"cloth128 jacobi 8-frame trajectory: worst rel err {worst:.3e}")


def _stats(x, v, inv_mass, h):
    """RMS violation of the DISTANCE rows and kinetic energy."""
    a, b, rest = h["a"], h["b"], h["rest"]
    dist = h["type"] == 0
    L = np.linalg.norm(x[a[dist]] - x[b[dist]], axis=1)
    rms = float(np.sqrt(np.mean((L - rest[dist]) ** 2)))
    m = np.where(inv_mass > 0, 1.0 / np.maximum(inv_mass, 1e-30), 0.0)
    ke = float(0.5 * np.sum(m[:, None] * v.astype(np.float64) ** 2))
    return rms, ke


@pytest.mark.parametrize("res", [48, 128])
def test_cloth_jacobi_600_frame_statistics(oracle, res):
    """Jacobi over 600 frames (48^2 cloth and config 2's 128^2 cloth, which
    lands on the floor): the 600-frame means of the RMS violation of the
    distance rows and of the kinetic energy agree with the oracle's within
    1 %, and so does the RMS violation over every 100-frame window.
    Instantaneous values are not compared: the swinging sheet is chaotic,
    and a one-ulp change of the oracle's own start positions moves its
    instantaneous kinetic energy by up to 10 % (while the 600-frame means
    move by 0.01 % and 0.4 %)."""
    from xpbd_amd import Solver
    from xpbd_amd.scenes import cloth_sheet

    s, _ = cloth_sheet(Solver, res=res, solve_mode="jacobi", device=DEV)
    h = s.host_arrays()
    orc = _oracle_scene(oracle, s)
    g, o = [], []
    for f in range(600):
        s.step()
        orc.step(1)
        g.append(_stats(s.positions(), s.velocities(), h["inv_mass"], h))
        o.append(_stats(orc.x, orc.v, h["inv_mass"], h))
    g, o = np.asarray(g), np.asarray(o)
    for w in range(6):
        rg, ro = g[100 * w:100 * (w + 1), 0].mean(), o[100 * w:100 * (w + 1), 0].mean()
        assert abs(rg - ro) <= 0.01 * ro, (w, rg, ro)
    (rg, kg), (ro, ko) = g.mean(axis=0), o.mean(axis=0)
    print(f"600-frame means: rms {rg:.4e} vs {ro:.4e}, ke {kg:.4e} vs {ko:.4e}")
    assert abs(rg - ro) <= 0.01 * ro, (rg, ro)
    assert abs(kg - ko) <= 0.01 * ko, (kg, ko)


@pytest.mark.parametrize("tiling", [(1, 0), (2, 1024), (1, 256)])
def test_jacobi_window_and_global_atomic_variants_agree(oracle, tiling):
    """The LDS-window kernel (windows that hold every endpoint, and windows
    that spill to global atomics) and the plain global-atomic kernel all match
    the oracle per substep."""
    from xpbd_amd import Solver
    from xpbd_amd.scenes import cloth_sheet

    def build(sub):
        kw = dict(res=64, solve_mode="jacobi", device=DEV)
        if sub == 1:
            kw.update(dt=float(np.float32(1.0 / 60.0 / 15.0)), substeps=1)
        s = cloth_sheet(Solver, **kw)[0]
        s.jacobi_kernel = "window"
        s._flush()
        s._tiling = tiling
        return s

    _substep_check(oracle, build, n_checks=3, frames_between=2, tol=1e-4)


@pytest.mark.parametrize("tile,reorder,iters", [(32, True, 2), (128, False, 3), (1024, True, 1)])
def test_jacobi_fused_tile_variants(oracle, tile, reorder, iters):
    """The fused tile kernel with tiny tiles (most rows are halo rows), no
    particle reordering, an odd iteration count (final state in the second
    ping-pong buffer) and the largest tile."""
    from xpbd_amd import Solver
    from xpbd_amd.scenes import cloth_sheet

    def build(sub):
        kw = dict(res=40, solve_mode="jacobi", device=DEV, iterations=iters)
        if sub == 1:
            kw.update(dt=float(np.float32(1.0 / 60.0 / 15.0)), substeps=1)
        s = cloth_sheet(Solver, **kw)[0]
        s.tile_size, s.reorder = tile, reorder
        return s

    _substep_check(oracle, build, n_checks=3, frames_between=2, tol=1e-4)


@pytest.mark.parametrize("tile", [64, 1024])
def test_jacobi_tile_is_bitwise_reproducible(tile):
    """The fused tile kernel sums corrections in exact fixed point, so two
    runs of the same scene agree bit for bit although the order of the
    atomic adds differs from run to run."""
    from xpbd_amd import Solver
    from xpbd_amd.scenes import cloth_sheet

    out = []
    for _ in range(2):
        s, _ = cloth_sheet(Solver, res=48, solve_mode="jacobi", device=DEV)
        s.tile_size = tile
        for _ in range(30):
            s.step()
        out.append((s.positions(), s.velocities(), s.lambdas()))
    for g, e in zip(out[0], out[1]):
        assert np.array_equal(g, e)


@pytest.mark.parametrize("gscale", [False, True])
def test_tile_first_equals_integrate_then_tile(gscale):
    """xpbd_jacobi_tile_first (integrate folded into the first iteration)
    gives bit for bit what xpbd_integrate_to followed by a first plain tile
    launch gives: positions, x_prev and multipliers."""
    import torch

    from xpbd_amd import Solver
    from xpbd_amd import _native as N
    from xpbd_amd.scenes import cloth_sheet

    s, _ = cloth_sheet(Solver, res=40, solve_mode="jacobi", device=DEV)
    s.tile_size = 64
    for _ in range(5):
        s.step()
    s._flush()
    lib = N.lib()
    n, m = s.num_bodies, s.num_constraints
    h = float(np.float32(s.dt / s.substeps))
    g = s.gravity
    gs = (torch.rand(n, device=DEV) + 0.5) if gscale else None
    ts, hs, hr = s._tile_dev
    X0, V0 = s.X.clone(), s.V.clone()
    # path A: integrate, then the plain first iteration
    XA, VA, XPA = torch.empty_like(X0), V0.clone(), torch.zeros_like(X0)
    XoA, LA = torch.empty_like(X0), torch.zeros(m, device=DEV)
    N.check(lib.xpbd_integrate_to(None, N.ptr(X0), N.ptr(XA), N.ptr(VA), N.ptr(XPA), *g, h,
                                  N.ptr(gs) if gs is not None else None, n))
    N.check(lib.xpbd_jacobi_tile(None, N.ptr(XA), N.ptr(XoA), N.ptr(XPA), N.ptr(VA), N.ptr(s.R),
                                 N.ptr(s.ANC), N.ptr(LA), N.ptr(LA), N.ptr(ts), N.ptr(hs),
                                 N.ptr(hr), h, 1, s._tile, n, s.jacobi_relax, 0, s.damping,
                                 s.max_speed, None))
    # path B: fused
    XoB, XPB, LB = torch.empty_like(X0), torch.zeros_like(X0), torch.zeros(m, device=DEV)
    N.check(lib.xpbd_jacobi_tile_first(None, N.ptr(X0), N.ptr(XoB), N.ptr(XPB), N.ptr(V0),
                                       N.ptr(s.R), N.ptr(s.ANC), N.ptr(LB), N.ptr(ts), N.ptr(hs),
                                       N.ptr(hr), h, *g, N.ptr(gs) if gs is not None else None,
                                       s._tile, n, s.jacobi_relax, None))
    torch.cuda.synchronize()
    assert torch.equal(XoA, XoB)
    assert torch.equal(XPA, XPB)
    assert torch.equal(LA, LB)
    assert torch.equal(V0, s.V)          # the fused launch leaves V alone


@pytest.mark.parametrize("gscale", [False, True])
def test_tile_next_equals_finalize_then_integrate(gscale):
    """xpbd_jacobi_tile_next (the last iteration of a substep with the next
    substep's integrate folded in) gives bit for bit what a finalizing
    xpbd_jacobi_tile launch followed by xpbd_integrate_to gives: predicted
    positions, x_prev and multipliers; V is left alone."""
    import torch

    from xpbd_amd import Solver
    from xpbd_amd import _native as N
    from xpbd_amd.scenes import cloth_sheet

    s, _ = cloth_sheet(Solver, res=40, solve_mode="jacobi", device=DEV)
    s.tile_size = 64
    s.max_speed = 0.5                      # the cap takes part
    for _ in range(12):
        s.step()
    s._flush()
    lib = N.lib()
    n, m = s.num_bodies, s.num_constraints
    h = float(np.float32(s.dt / s.substeps))
    g = s.gravity
    gs = (torch.rand(n, device=DEV) + 0.5) if gscale else None
    ts, hs, hr = s._tile_dev
    # a mid-substep state: integrate, one plain iteration
    X0, V0, XP0 = torch.empty_like(s.X), s.V.clone(), torch.zeros_like(s.X)
    L0 = torch.zeros(m, device=DEV)
    N.check(lib.xpbd_integrate_to(None, N.ptr(s.X), N.ptr(X0), N.ptr(V0), N.ptr(XP0), *g, h,
                                  N.ptr(gs) if gs is not None else None, n))
    X1 = torch.empty_like(X0)
    N.check(lib.xpbd_jacobi_tile(None, N.ptr(X0), N.ptr(X1), N.ptr(XP0), N.ptr(V0), N.ptr(s.R),
                                 N.ptr(s.ANC), N.ptr(L0), N.ptr(L0), N.ptr(ts), N.ptr(hs),
                                 N.ptr(hr), h, 1, s._tile, n, s.jacobi_relax, 0, s.damping,
                                 s.max_speed, None))
    # path A: finalizing iteration, then integrate
    XA, VA, XPA, LA = torch.empty_like(X0), V0.clone(), XP0.clone(), torch.zeros(m, device=DEV)
    N.check(lib.xpbd_jacobi_tile(None, N.ptr(X1), N.ptr(XA), N.ptr(XPA), N.ptr(VA), N.ptr(s.R),
                                 N.ptr(s.ANC), N.ptr(L0), N.ptr(LA), N.ptr(ts), N.ptr(hs),
                                 N.ptr(hr), h, 0, s._tile, n, s.jacobi_relax, 1, s.damping,
                                 s.max_speed, None))
    N.check(lib.xpbd_integrate_to(None, N.ptr(XA), N.ptr(XA), N.ptr(VA), N.ptr(XPA), *g, h,
                                  N.ptr(gs) if gs is not None else None, n))
    # path B: one launch
    XB, XPB, LB = torch.empty_like(X0), XP0.clone(), torch.zeros(m, device=DEV)
    VB = V0.clone()
    N.check(lib.xpbd_jacobi_tile_next(None, N.ptr(X1), N.ptr(XB), N.ptr(XPB), N.ptr(s.R),
                                      N.ptr(s.ANC), N.ptr(L0), N.ptr(LB), N.ptr(ts), N.ptr(hs),
                                      N.ptr(hr), h, 0, *g,
                                      N.ptr(gs) if gs is not None else None, s._tile, n,
                                      s.jacobi_relax, s.damping, s.max_speed, None))
    torch.cuda.synchronize()
    assert torch.equal(XA, XB)
    assert torch.equal(XPA, XPB)
    assert torch.equal(LA, LB)
    assert torch.equal(VB, V0)
    with pytest.raises(RuntimeError):       # Xin == Xout is refused
        N.check(lib.xpbd_jacobi_tile_next(None, N.ptr(X1), N.ptr(X1), N.ptr(XPB), N.ptr(s.R),
                                          None, N.ptr(L0), N.ptr(LB), N.ptr(ts), N.ptr(hs),
                                          N.ptr(hr), h, 0, *g, None, s._tile, n,
                                          s.jacobi_relax, s.damping, s.max_speed, None))


@pytest.mark.parametrize("iters", [1, 2, 3])
def test_frame_plan_integrate_folding_is_exact(iters):
    """The three ways of running integrate in a tile frame (a separate pass,
    folded into each substep's first launch, folded into the previous
    substep's last launch) give identical positions, velocities and
    multipliers, with per-particle gravity scales and ATTACH rows."""
    from xpbd_amd import Solver
    from xpbd_amd.scenes import cloth_sheet
    from xpbd_amd.solver import Body

    out = {}
    for fuse in (False, "first", True):
        s, _ = cloth_sheet(Solver, res=36, solve_mode="jacobi", device=DEV, iterations=iters)
        s.tile_size = 64
        s.block_frames = False
        s.fuse_first = fuse
        s.set_gravity_scale(np.linspace(0.5, 1.5, s.num_bodies))
        s.add_attach(Body(5), (0.1, 0.6, 0.0), compliance=1e-6)
        s.step(8)
        out[str(fuse)] = (s.positions(), s.velocities(), s.lambdas())
    for k in ("first", "True"):
        for g, e in zip(out[k], out["False"]):
            assert np.array_equal(g, e), (k, iters)


@pytest.mark.parametrize("substeps,iters", [(1, 1), (1, 2), (2, 1), (3, 3)])
def test_integrate_folding_edge_shapes(substeps, iters):
    """The integrate folds on the edge shapes of a frame (one substep: no
    next-substep fold at all; one iteration: the first substep integrates in
    place and every launch finalizes) on a ragged scene (49 particles, one
    partial 64-particle tile, 25 extra particles in a second one at tile 32):
    identical to the separate integrate pass."""
    from xpbd_amd import Solver
    from xpbd_amd.scenes import cloth_sheet

    for tile in (64, 32):
        out = {}
        for fuse in (False, "first", True):
            s, _ = cloth_sheet(Solver, res=7, solve_mode="jacobi", device=DEV,
                               substeps=substeps, iterations=iters)
            s.tile_size = tile
            s.block_frames = False
            s.fuse_first = fuse
            s.step(12)
            out[str(fuse)] = (s.positions(), s.velocities(), s.lambdas())
        for k in ("first", "True"):
            for g, e in zip(out[k], out["False"]):
                assert np.array_equal(g, e), (k, tile, substeps, iters)


def test_autotune_restores_state_and_keeps_results():
    """The autotuner times every launch shape without advancing the scene,
    and the shape it picks steps to the same state (to the Jacobi parity
    bar) as the default shape."""
    from xpbd_amd import Solver, autotune
    from xpbd_amd.scenes import cloth_sheet

    s, _ = cloth_sheet(Solver, res=40, solve_mode="jacobi", device=DEV)
    ref, _ = cloth_sheet(Solver, res=40, solve_mode="jacobi", device=DEV)
    for _ in range(3):
        s.step()
        ref.step()
    x0, v0 = s.positions(), s.velocities()
    best, times = autotune.tune(s, frames=2)
    assert best in times and len(times) >= 4
    assert np.array_equal(s.positions(), x0) and np.array_equal(s.velocities(), v0)
    for _ in range(5):
        s.step()
        ref.step()
    assert _rel(s.positions(), ref.positions()) < 1e-4
    assert _rel(s.velocities(), ref.velocities()) < 2e-3
    # cached per scene signature
    again, t2 = autotune.tune(s, frames=2)
    assert again == best and t2 == {}


@pytest.mark.parametrize("scene,mode", [("chain", "gs"), ("cloth", "gs"), ("cloth", "jacobi")])
def test_block_frames_equal_launch_path(scene, mode):
    """Whole frames in one workgroup (small scenes) give bit for bit what
    the per-phase launches give: GS (the same row arithmetic in the same
    color order) and Jacobi (the fused tile kernel's arithmetic with exact
    sums)."""
    from xpbd_amd import Solver
    from xpbd_amd.scenes import cloth_sheet, hanging_chain

    out = []
    for block in (True, False):
        if scene == "chain":
            s = hanging_chain(Solver, device=DEV)
        else:
            s = cloth_sheet(Solver, res=24, solve_mode=mode, device=DEV)[0]
        s.block_frames = block
        s.step(5)
        s.step()
        assert s._block_active == block
        out.append((s.positions(), s.velocities(), s.lambdas()))
    for a, b in zip(*out):
        assert np.array_equal(a, b)


def test_block_frames_capacity_falls_back():
    """Above the workgroup's LDS capacity the solver keeps the launch path."""
    from xpbd_amd import Solver
    from xpbd_amd import _native as N
    from xpbd_amd.scenes import cloth_sheet

    cap = int(N.lib().xpbd_frame_block_capacity(0))
    s = cloth_sheet(Solver, res=int(np.ceil(np.sqrt(cap + 1))), solve_mode="jacobi",
                    device=DEV)[0]
    s.step()
    assert not s._block_active and s.num_bodies > cap


def test_graph_replay_equals_eager_launches():
    from xpbd_amd import Solver
    from xpbd_amd.scenes import cloth_sheet

    out = []
    for use_graph in (True, False):
        s, _ = cloth_sheet(Solver, res=32, solve_mode="gs", device=DEV)
        s.use_graph = use_graph
        s.block_frames = False       # the per-phase launch path, graph or eager
        for _ in range(20):
            s.step()
        out.append((s.positions(), s.velocities(), s.lambdas()))
    for g, e in zip(out[0], out[1]):
        assert (g == e).all()


def _mixed_scene(mode, substeps=8, dt=1 / 60):
    from xpbd_amd import Body, Solver

    s = Solver(dt=dt, substeps=substeps, iterations=3, solve_mode=mode, device=DEV,
               max_speed=2.0, damping=0.01)
    rng = np.random.default_rng(7)
    for i in range(50):
        s.add_particle(tuple(rng.uniform(-1, 1, 3) + [0, 1.2, 0]), mass=rng.uniform(0.5, 2),
                       velocity=tuple(rng.normal(size=3)), collide=False)
    for i in range(49):
        s.add_distance(Body(i), Body(i + 1), compliance=float(rng.choice([0.0, 1e-5])))
    for i in range(0, 50, 5):
        s.add_attach(Body(i), tuple(rng.uniform(-1, 1, 3)), compliance=1e-4, rest=0.1)
    for i in range(50):
        s.add_floor_contact(Body(i), floor_y=0.3, friction=0.5)
    return s


def test_attach_and_mixed_rows_match_oracle(oracle):
    """ATTACH + DISTANCE + FLOOR rows with friction and the velocity cap
    active: GS frame by frame, Jacobi per substep."""
    _compare_frames(oracle, _mixed_scene("gs"), 30, 1e-5)

    def build(sub):
        if sub == 1:
            return _mixed_scene("jacobi", substeps=1, dt=float(np.float32(1 / 60 / 8)))
        return _mixed_scene("jacobi")

    _substep_check(oracle, build, n_checks=5, frames_between=4, tol=1e-4)


def test_edge_cases():
    from xpbd_amd import Body, Solver

    s = Solver(device=DEV)
    s.step()                                   # empty scene: no-op
    s.add_particle((0, 1, 0), mass=1.0)
    s.step()                                   # particles, no rows
    assert s.positions()[0, 1] < 1.0
    s = Solver(device=DEV, iterations=0)
    a = s.add_particle((0, 1, 0), mass=0.0)
    b = s.add_particle((1, 1, 0), mass=1.0)
    s.add_distance(a, b)
    s.step()                                   # iterations = 0: free fall only
    assert s.positions()[1, 1] < 1.0
    s = Solver(device=DEV, max_speed=0.0)      # cap disabled
    s.add_particle((0, 0, 0), mass=1.0, velocity=(100.0, 0, 0))
    s.step()
    assert s.velocities()[0, 0] == pytest.approx(100.0)
    # knob change re-captures; live compliance update is in place
    s = Solver(device=DEV, solve_mode="gs", substeps=5)
    a = s.add_particle((0, 2, 0), mass=0.0)
    b = s.add_particle((0, 1, 0), mass=1.0)
    s.add_distance(a, b, rest=1.0, compliance=1e-2)
    s.step()
    ptr = s.R.data_ptr()
    s.set_all_compliance(0.0)
    assert s.R.data_ptr() == ptr
    s.substeps = 10
    for _ in range(60):
        s.step()
    assert abs(np.linalg.norm(s.positions()[1] - s.positions()[0]) - 1.0) < 1e-3
    del Body


# ---- the reference's invariant tests (tests/test_solver.py), on the GPU ----
def test_free_fall_matches_analytic():
    from xpbd_amd import Solver

    s = Solver(dt=1 / 60, substeps=10, iterations=1, device=DEV)
    b = s.add_particle((0, 0, 0), mass=1.0, collide=False)
    for _ in range(60):
        s.step()
    assert abs(s.velocities()[b.index, 1] - (-9.81)) < 0.05
    assert abs(s.positions()[b.index, 1] - 0.5 * -9.81) < 0.05


def test_static_body_never_moves():
    from xpbd_amd import Solver

    s = Solver(dt=1 / 60, substeps=10, device=DEV)
    a = s.add_particle((0, 1, 0), mass=0.0, collide=False)
    for _ in range(30):
        s.step()
    assert np.allclose(s.positions()[a.index], [0, 1, 0])


@pytest.mark.parametrize("mode", ["gs", "jacobi"])
def test_hard_distance_holds_length(mode):
    from xpbd_amd import Solver

    s = Solver(dt=1 / 60, substeps=15, iterations=2, device=DEV, solve_mode=mode)
    a = s.add_particle((0, 2, 0), mass=0.0, collide=False)
    b = s.add_particle((0.5, 2, 0), mass=1.0, collide=False)
    s.add_distance(a, b, rest=0.5, compliance=0.0)
    for _ in range(120):
        s.step()
    L = np.linalg.norm(s.positions()[b.index] - s.positions()[a.index])
    assert abs(L - 0.5) < 1e-3


@pytest.mark.parametrize("iters", [1, 4, 16])
def test_compliance_iteration_independent(iters):
    from xpbd_amd import Solver

    alpha = 1e-3
    s = Solver(dt=1 / 60, substeps=20, iterations=iters, device=DEV, solve_mode="gs",
               damping=0.05)
    a = s.add_particle((0, 2, 0), mass=0.0, collide=False)
    b = s.add_particle((0, 1, 0), mass=1.0, collide=False)
    s.add_distance(a, b, rest=1.0, compliance=alpha)
    for _ in range(800):
        s.step()
    stretch = np.linalg.norm(s.positions()[b.index] - s.positions()[a.index]) - 1.0
    assert abs(stretch - 9.81 * alpha) < 0.1 * 9.81 * alpha


def test_floor_contact_rest():
    from xpbd_amd import Shape, Solver

    s = Solver(dt=1 / 60, substeps=10, iterations=2, device=DEV)
    b = s.add_particle((0, 2, 0), mass=1.0, shape=Shape("sphere", (0.1,)), collide=False)
    s.add_floor_contact(b, floor_y=0.1, friction=0.5)
    for _ in range(180):
        s.step()
    assert abs(s.positions()[b.index, 1] - 0.1) < 1e-3
    assert abs(s.velocities()[b.index, 1]) < 1e-2


@pytest.mark.parametrize("mode", ["gs", "jacobi"])
def test_two_spheres_do_not_overlap(mode):
    from xpbd_amd import Shape, Solver

    s = Solver(dt=1 / 60, substeps=15, iterations=2, device=DEV, solve_mode=mode)
    a = s.add_particle((0.0, 1.0, 0), mass=1.0, shape=Shape("sphere", (0.2,)))
    b = s.add_particle((0.15, 0.4, 0), mass=1.0, shape=Shape("sphere", (0.2,)))
    s.add_floor_contact(a, floor_y=0.2, friction=0.4)
    s.add_floor_contact(b, floor_y=0.2, friction=0.4)
    s.enable_self_collision(True, default_friction=0.4)
    for _ in range(240):
        s.step()
    d = np.linalg.norm(s.positions()[a.index] - s.positions()[b.index])
    assert d >= 0.4 - 2e-3


@pytest.mark.parametrize("mode", ["gs", "jacobi"])
def test_stack_column_settles(mode):
    from xpbd_amd import Shape, Solver

    s = Solver(dt=1 / 60, substeps=20, iterations=2, device=DEV, solve_mode=mode, max_speed=0.0)
    r = 0.16
    for k in range(4):
        b = s.add_particle((0, r + k * 2.1 * r, 0), mass=1.0, shape=Shape("sphere", (r,)),
                           friction=0.3)
        s.add_floor_contact(b, floor_y=r, friction=0.3)
    s.enable_self_collision(True, margin=0.05, default_friction=0.3)
    for _ in range(220):
        s.step()
    ys = np.sort(s.positions()[:, 1])
    assert ys.max() < 1.5
    assert np.all(np.isfinite(s.positions()))
    assert np.allclose(np.diff(ys), 2 * r, atol=5e-3)


def test_friction_decelerates_slide():
    from xpbd_amd import Shape, Solver

    s = Solver(dt=1 / 60, substeps=10, iterations=2, device=DEV)
    b = s.add_particle((0, 0.1, 0), mass=1.0, shape=Shape("sphere", (0.1,)), collide=False)
    s.add_floor_contact(b, floor_y=0.1, friction=0.5)
    s.set_velocity(b, (3.0, 0, 0))
    for _ in range(120):
        s.step()
    assert abs(s.velocities()[b.index, 0]) < 0.5


def test_nan_position_is_sanitized():
    """A NaN position is reset to its initial value before the broad phase
    (solver.py:469-483), forced from the host."""
    from xpbd_amd import Shape, Solver

    s = Solver(dt=1 / 60, substeps=4, device=DEV)
    a = s.add_particle((0, 1, 0), mass=1.0, shape=Shape("sphere", (0.1,)))
    s.add_particle((1, 1, 0), mass=1.0, shape=Shape("sphere", (0.1,)))
    s.enable_self_collision(True)
    s.step()
    s.set_position(a, (float("nan"), 0.0, 0.0))
    s.step()
    assert np.all(np.isfinite(s.positions()))


@pytest.mark.parametrize("mode", ["jacobi", "gs"])
def test_scene_file_resumes_the_trajectory(tmp_path, mode):
    """A cloth stepped 10 frames, saved (xpbd_amd.scene_io) and loaded into a
    new solver continues exactly as the original: 5 more frames each,
    positions and velocities bit for bit (Jacobi sums are exact, so the new
    solver's own particle order does not matter)."""
    from xpbd_amd import Solver, scene_io
    from xpbd_amd.scenes import cloth_sheet

    s = cloth_sheet(Solver, res=32, solve_mode=mode, device=DEV)[0]
    s.step(10)
    p = tmp_path / "cloth.npz"
    scene_io.save(s, p)
    t = scene_io.load(p, device=DEV)
    s.step(5)
    t.step(5)
    assert np.array_equal(s.positions(), t.positions())
    assert np.array_equal(s.velocities(), t.velocities())


def test_batched_world_files_resume():
    """The batched world saved to scene files and loaded back steps on: same
    particle and body counts, finite states after a frame, positions close
    to the original's after the same frame (the rigid part sums with float
    atomics)."""
    import tempfile

    from xpbd_amd.batch import BatchedWorld, draw

    w = BatchedWorld(draw(32, 3), device=DEV)
    w.step()
    with tempfile.TemporaryDirectory() as d:
        w.save(d + "/b")
        v = BatchedWorld.load(d + "/b", device=DEV)
    assert v.particles.num_bodies == w.particles.num_bodies
    assert (v.rigid is None) == (w.rigid is None)
    w.step()
    v.step()
    x0, x1 = w.particles.positions(), v.particles.positions()
    assert np.all(np.isfinite(x1))
    assert np.abs(x0 - x1).max() <= 1e-4
    if w.rigid is not None:
        assert np.abs(w.rigid.positions() - v.rigid.positions()).max() <= 1e-3


def test_batched_world_step_equals_its_solvers_stepped_alone():
    """The world's step (particle frame on the main stream, rigid frame on a
    second stream, overlapping) gives the states that stepping its two
    solvers one after the other on one stream gives, to float-atomic
    rounding."""
    from xpbd_amd.batch import BatchedWorld, draw

    spec = draw(40, 5, mix={"cloth": 0.4, "pile": 0.2, "chain": 0.2, "boxes": 0.2})
    assert any(sc.kind == "boxes" for sc in spec.scenes)
    w = BatchedWorld(spec, device=DEV)
    v = BatchedWorld(spec, device=DEV)
    for _ in range(6):
        w.step()
        v.particles.step()
        v.rigid.step()
    xp_w, xp_v = w.particles.positions(), v.particles.positions()
    assert np.all(np.isfinite(xp_w))
    assert np.abs(xp_w - xp_v).max() <= 1e-4
    assert np.abs(w.rigid.positions() - v.rigid.positions()).max() <= 1e-3


def _sphere_pile(k=343, seed=5, **kw):
    from xpbd_amd import Solver

    s = Solver(dt=1 / 60, substeps=10, iterations=2, solve_mode="jacobi", device=DEV, **kw)
    rng = np.random.default_rng(seed)
    side = int(round(k ** (1 / 3)))
    g = np.arange(k)
    ix, rem = g % side, g // side
    iz, iy = rem % side, rem // side
    pos = np.stack([ix * 0.102, 0.06 + iy * 0.102, iz * 0.102], 1)
    pos += rng.uniform(-3e-3, 3e-3, pos.shape)
    s.add_particles(pos, np.ones(k), radius=0.05, collide=True, friction=0.5)
    s.add_floor_contacts(np.arange(k), floor_y=0.05)
    return s


@pytest.mark.parametrize("moving", [False, True])
def test_device_contacts_match_host_rows_in_order(moving):
    """Device CONTACT generation (xpbd_sphere_contacts) gives the host broad
    phase's rows (the reference's solver.py:275-342 with the enumeration
    order of _grid_candidate_pairs) in the same order: same (a, b) in API
    indices, same rest lengths and friction bit for bit, pairs joined by a
    DISTANCE row skipped; on a jittered sphere pile in the Morton device
    layout, at rest and with random velocities (larger search margin)."""
    from xpbd_amd import Body

    def build():
        s = _sphere_pile()
        for i in range(0, 60, 3):        # chain a few neighbours: those pairs are skipped
            s.add_distance(Body(i), Body(i + 1), compliance=1e-6)
        if moving:
            rng = np.random.default_rng(2)
            s._v.set(rng.normal(0, 0.4, (s.num_bodies, 3)))
        return s

    d = build()
    d.enable_device_contacts(True, margin=0.01)
    d._flush()
    assert d._pperm is not None                   # the device layout is permuted
    d._device_contact_pass()
    cnt = d.n_contacts
    rows = d.CROWS[:cnt].cpu().numpy()
    perm = d._pperm
    a = perm[(rows[:, 0] & 0x0FFFFFFF).astype(np.int64)]
    b = perm[rows[:, 1].astype(np.int64)]
    h = build()
    h.enable_self_collision(True, margin=0.01)
    h._flush()                                    # velocities on the device (the margin)
    h._rebuild_contacts()
    sel = h._rt.array() == 3
    ha, hb = h._ra.array()[sel], h._rb.array()[sel]
    assert cnt == len(ha) > 300
    assert np.array_equal(a, ha) and np.array_equal(b, hb)
    assert np.array_equal(rows[:, 2], h._rrest.array()[sel].astype(np.float32).view(np.int32))
    assert np.array_equal(rows[:, 3], h._rmu.array()[sel].astype(np.float32).view(np.int32))
    joined = {(i, i + 1) for i in range(0, 60, 3)}
    assert not joined & {(min(x, y), max(x, y)) for x, y in zip(a, b)}


def test_device_contacts_fused_integrate_matches_separate():
    """With device-generated CONTACT rows the first iteration of a substep
    folds integrate in (xpbd_solve_jacobi_dyn_first + xpbd_jacobi_tile_first
    with the ACC segment); against the separate integrate launch the states
    agree to float-atomic rounding after 5 frames and the pile rests on the
    floor without sinking through."""
    out = []
    for fuse in (True, False):
        s = _sphere_pile()
        s.enable_device_contacts(True, margin=0.01)
        s.fuse_first = fuse
        s.block_frames = False
        s.step(5)
        out.append((s.positions(), s.velocities()))
    (x0, v0), (x1, v1) = out
    assert np.abs(x0 - x1).max() <= 1e-4
    assert np.abs(v0 - v1).max() <= 5e-3
    assert x0[:, 1].min() > 0.05 - 0.01

