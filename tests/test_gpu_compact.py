"""The fused tile launches over the compact tables (xpbd_jacobi_tile*_cx,
xpbd_tile_compact) against the same launches over the int4 rows and halo
list: positions, x_prev, velocities and multipliers bit for bit, with and
without per-particle floors; and whole Solver plans with compact rows
against plans without them (compact_rows = False)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _scene(res=40, land=30):
    """Cloth with ATTACH rows, a particle with two FLOOR rows (read whole)
    and, after ``land`` frames, FLOOR rows active with friction."""
    from xpbd_amd import Body, Solver
    from xpbd_amd.scenes import cloth_sheet

    s, _ = cloth_sheet(Solver, res=res, solve_mode="jacobi", device=DEV)
    x = s.host_arrays()["x"]
    for i in range(0, res * res, 53):    # soft tethers to points just above the floor
        s.add_attach(Body(i), (x[i, 0], 0.02, x[i, 2]), compliance=1e-3, rest=0.05)
    s.add_floor_contact(Body(7), floor_y=0.01, friction=0.3)
    s.tile_size = 64
    s.max_speed = 0.7
    for _ in range(land):
        s.step()
    s._flush()
    return s


def _tables(s, pf):
    from xpbd_amd import _native as N

    c = s._cx_dev
    cs = N.TileCompact()
    for k in ("dist_end", "gen_end", "ci", "cc", "halo2"):
        setattr(cs, k, N.ptr(c[k]))
    if pf:
        cs.pf, cs.pflam = N.ptr(c["pf"]), N.ptr(c["pflam"])
    cs.general = 1            # ATTACH rows, the doubled FLOOR row
    return cs


def _check_lam(s, L_old, L_new, pf, pflam):
    import torch

    if not pf:
        assert torch.equal(L_old, L_new)
        return
    rows = torch.as_tensor(s._cx["pf_rows"], device=DEV)
    part = torch.as_tensor(s._cx["pf_part"], device=DEV)
    keep = torch.ones_like(L_old, dtype=torch.bool)
    keep[rows] = False
    assert torch.equal(L_old[keep], L_new[keep])
    assert torch.equal(L_old[rows], pflam[part])


@pytest.mark.parametrize("pf", [True, False])
@pytest.mark.parametrize("first,finalize", [(1, 0), (0, 0), (0, 1), (1, 1)])
def test_plain_cx_launch_equals_int4_launch(pf, first, finalize):
    import torch

    from xpbd_amd import _native as N

    s = _scene()
    lib = N.lib()
    n, m = s.num_bodies, s.num_constraints
    h = float(np.float32(s.dt / s.substeps))
    ts, hs, hr = s._tile_dev
    X0 = s.X.clone()
    X0[::3, 1] = -0.002          # a third of the particles below their floor
    XP = s.X.clone()
    Lin = torch.rand(m, device=DEV) * 1e-3
    cs = _tables(s, pf)
    if pf:   # per-particle multipliers start where the int4 launch reads them
        s._cx_dev["pflam"][torch.as_tensor(s._cx["pf_part"], device=DEV)] = \
            Lin[torch.as_tensor(s._cx["pf_rows"], device=DEV)]
    VA, VB = s.V.clone(), s.V.clone()
    XA, XB = torch.empty_like(X0), torch.empty_like(X0)
    LA, LB = torch.zeros(m, device=DEV), torch.zeros(m, device=DEV)
    N.check(lib.xpbd_jacobi_tile(None, N.ptr(X0), N.ptr(XA), N.ptr(XP), N.ptr(VA), N.ptr(s.R),
                                 N.ptr(s.ANC), N.ptr(Lin), N.ptr(LA), N.ptr(ts), N.ptr(hs),
                                 N.ptr(hr), h, first, s._tile, n, s.jacobi_relax, finalize,
                                 s.damping, s.max_speed, None))
    N.check(lib.xpbd_jacobi_tile_cx(None, N.ptr(X0), N.ptr(XB), N.ptr(XP), N.ptr(VB),
                                    N.ptr(s.R), N.ptr(s.ANC), N.ptr(Lin), N.ptr(LB), N.ptr(ts),
                                    N.ptr(hs), N.ctypes.addressof(cs), h, first, s._tile, n,
                                    s.jacobi_relax, finalize, s.damping, s.max_speed, None))
    torch.cuda.synchronize()
    assert torch.equal(XA, XB)
    assert torch.equal(VA, VB)
    _check_lam(s, LA, LB, pf, s._cx_dev["pflam"])
    if pf:   # the floors were active: the per-particle path did work
        rows = s._cx["pf_rows"]
        assert int(torch.count_nonzero(LA[torch.as_tensor(rows, device=DEV)])) > 0


@pytest.mark.parametrize("pf", [True, False])
def test_first_and_next_cx_launches_equal_int4_launches(pf):
    import torch

    from xpbd_amd import _native as N

    s = _scene()
    lib = N.lib()
    n, m = s.num_bodies, s.num_constraints
    h = float(np.float32(s.dt / s.substeps))
    g = s.gravity
    gs = torch.rand(n, device=DEV) + 0.5
    ts, hs, hr = s._tile_dev
    cs = _tables(s, pf)
    X0 = s.X.clone()
    X0[::3, 1] = -0.002
    V0 = s.V.clone()
    # first: integrate folded in
    XA, XB = torch.empty_like(X0), torch.empty_like(X0)
    PA, PB = torch.zeros_like(X0), torch.zeros_like(X0)
    LA, LB = torch.zeros(m, device=DEV), torch.zeros(m, device=DEV)
    s._cx_dev["pflam"].fill_(123.0)       # a zero-start launch must reset them
    V0[::3, 1] = -0.3                      # predicted below the floor too
    N.check(lib.xpbd_jacobi_tile_first(None, N.ptr(X0), N.ptr(XA), N.ptr(PA), N.ptr(V0),
                                       N.ptr(s.R), N.ptr(s.ANC), N.ptr(LA), N.ptr(ts), N.ptr(hs),
                                       N.ptr(hr), h, *g, N.ptr(gs), s._tile, n, s.jacobi_relax,
                                       None))
    N.check(lib.xpbd_jacobi_tile_first_cx(None, N.ptr(X0), N.ptr(XB), N.ptr(PB), N.ptr(V0),
                                          N.ptr(s.R), N.ptr(s.ANC), N.ptr(LB), N.ptr(ts),
                                          N.ptr(hs), N.ctypes.addressof(cs), h, *g, N.ptr(gs),
                                          s._tile, n, s.jacobi_relax, None))
    torch.cuda.synchronize()
    assert torch.equal(XA, XB) and torch.equal(PA, PB)
    if pf:   # particles with a floor: reset or set; others untouched
        has = torch.isfinite(s._cx_dev["pf"][:, 0])
        assert not torch.any(s._cx_dev["pflam"][has] == 123.0)
    _check_lam(s, LA, LB, pf, s._cx_dev["pflam"])
    # next: the last iteration with the next substep's integrate
    YA, YB = torch.empty_like(X0), torch.empty_like(X0)
    QA, QB = PA.clone(), PB.clone()
    MA, MB = torch.zeros(m, device=DEV), torch.zeros(m, device=DEV)
    N.check(lib.xpbd_jacobi_tile_next(None, N.ptr(XA), N.ptr(YA), N.ptr(QA), N.ptr(s.R),
                                      N.ptr(s.ANC), N.ptr(LA), N.ptr(MA), N.ptr(ts), N.ptr(hs),
                                      N.ptr(hr), h, 0, *g, N.ptr(gs), s._tile, n,
                                      s.jacobi_relax, s.damping, s.max_speed, None))
    N.check(lib.xpbd_jacobi_tile_next_cx(None, N.ptr(XB), N.ptr(YB), N.ptr(QB), N.ptr(s.R),
                                         N.ptr(s.ANC), N.ptr(LB), N.ptr(MB), N.ptr(ts),
                                         N.ptr(hs), N.ctypes.addressof(cs), h, 0, *g, N.ptr(gs),
                                         s._tile, n, s.jacobi_relax, s.damping, s.max_speed,
                                         None))
    torch.cuda.synchronize()
    assert torch.equal(YA, YB) and torch.equal(QA, QB)
    _check_lam(s, MA, MB, pf, s._cx_dev["pflam"])


def _pair(build, frames):
    """The same scene stepped with and without compact rows."""
    a, b = build(), build()
    b.compact_rows = False
    for _ in range(frames):
        a.step()
        b.step()
    return a, b


def _equal(a, b):
    assert np.array_equal(a.positions(), b.positions())
    assert np.array_equal(a.velocities(), b.velocities())
    assert np.array_equal(a.lambdas(), b.lambdas())


@pytest.mark.parametrize("iters,tile", [(2, 64), (3, 128), (1, 64), (2, 1024)])
def test_solver_compact_plan_equals_int4_plan(iters, tile):
    """Cloth landing on its floor with friction (per-particle floors active),
    ATTACH rows and a doubled FLOOR row: frames with compact rows equal
    frames without, positions, velocities and multipliers."""
    from xpbd_amd import Body, Solver
    from xpbd_amd.scenes import cloth_sheet

    def build():
        s, _ = cloth_sheet(Solver, res=48, solve_mode="jacobi", device=DEV, iterations=iters)
        x = s.host_arrays()["x"]
        for i in range(0, 48 * 48, 97):
            s.add_attach(Body(i), (x[i, 0], 0.02, x[i, 2]), compliance=1e-3, rest=0.05)
        s.add_floor_contact(Body(9), floor_y=0.02, friction=0.2)
        s.tile_size = tile
        s.tile_pair = False
        return s

    a, b = _pair(build, 45)
    assert a._plan_pf and not b._plan_pf
    _equal(a, b)
    assert np.count_nonzero(a.lambdas()[a._rt.array() == 2]) > 0   # floors were active


def test_solver_compact_plan_with_self_collision_and_pair_launch():
    """CONTACT rows rebuilt on the host (read whole, integrate in the first
    launch) and the two-iteration pair launch (floors stay in the rows):
    both equal their int4 plans."""
    from xpbd_amd import Body, Solver
    from xpbd_amd.scenes import cloth_sheet

    def stack():
        s = Solver(dt=1 / 60, substeps=10, iterations=2, solve_mode="jacobi", device=DEV)
        rng = np.random.default_rng(3)
        for i in range(120):
            p = (0.11 * (i % 5), 0.06 + 0.11 * (i // 25), 0.11 * ((i // 5) % 5))
            s.add_particle(tuple(np.add(p, rng.normal(0, 1e-3, 3))), mass=1.0, radius=0.05)
        for i in range(120):
            s.add_floor_contact(Body(i), friction=0.4)
        s.enable_self_collision(True)
        return s

    a, b = _pair(stack, 20)
    _equal(a, b)

    def cloth():
        s, _ = cloth_sheet(Solver, res=32, solve_mode="jacobi", device=DEV, iterations=2)
        s.tile_pair = True
        return s

    a, b = _pair(cloth, 40)
    assert a._pair_dev is not None and not a._plan_pf
    _equal(a, b)


def test_live_knobs_reach_the_compact_tables():
    """set_all_compliance / set_all_friction rewrite the compact copies in
    place: the next frames still equal the int4 plan's."""
    from xpbd_amd import Solver
    from xpbd_amd.scenes import cloth_sheet

    def build():
        s, _ = cloth_sheet(Solver, res=40, solve_mode="jacobi", device=DEV)
        s.tile_pair = False
        return s

    a, b = _pair(build, 35)
    for s in (a, b):
        s.set_all_compliance(1e-5)
        s.set_all_friction(0.05)
    for _ in range(10):
        a.step()
        b.step()
    _equal(a, b)


@pytest.mark.parametrize("predict", [0, 1])
@pytest.mark.parametrize("general", [True, False])
@pytest.mark.parametrize("pf", [True, False])
@pytest.mark.parametrize("on", [1, 0])
def test_warm_cx_launch_equals_int4_launch(predict, general, pf, on):
    """xpbd_jacobi_tile_warm_cx gives the int4 launch's lambda0 and
    positions bit for bit: a cloth with ATTACH rows and a doubled FLOOR row
    (rows read whole), or a bare cloth; with per-particle floors the single
    FLOOR rows' lambda0 lands in pflam (zeros when the start is off); the
    network output is random."""
    import torch

    from xpbd_amd import _native as N
    from xpbd_amd import Solver
    from xpbd_amd.scenes import cloth_sheet

    if general:
        s = _scene()
    else:
        s, _ = cloth_sheet(Solver, res=40, solve_mode="jacobi", device=DEV)
        s.tile_size = 64
        for _ in range(30):
            s.step()
        s._flush()
    lib = N.lib()
    n, m = s.num_bodies, s.num_constraints
    h = float(np.float32(s.dt / s.substeps))
    ts, hs, hr = s._tile_dev
    cs = _tables(s, pf)
    cs.general = int(general or not pf)
    X0 = s.X.clone()
    X0[::3, 1] = -0.002
    Y = torch.randn(m, device=DEV).to(torch.bfloat16)
    ws_on = torch.full((1,), on, dtype=torch.int32, device=DEV)
    s._cx_dev["pflam"].fill_(123.0)
    outs = []
    for fn, tab in ((lib.xpbd_jacobi_tile_warm, N.ptr(hr)),
                    (lib.xpbd_jacobi_tile_warm_cx, N.ctypes.addressof(cs))):
        Xo, XPo, Lo = torch.empty_like(X0), s.XP.clone(), torch.zeros(m, device=DEV)
        N.check(fn(None, N.ptr(X0), N.ptr(Xo), N.ptr(XPo), N.ptr(s.V), N.ptr(s.R), N.ptr(s.ANC),
                   N.ptr(Lo), N.ptr(ts), N.ptr(hs), tab, N.ptr(Y), h, *s.gravity, None, s._tile,
                   n, s.jacobi_relax, predict, N.ptr(ws_on), 0))
        outs.append((Xo, XPo, Lo))
    torch.cuda.synchronize()
    (xa, pa, la), (xb, pb, lb) = outs
    assert torch.equal(xa, xb) and torch.equal(pa, pb)
    if pf:
        _check_lam(s, la, lb, True, s._cx_dev["pflam"])
    else:
        assert torch.equal(la, lb)
        assert bool(torch.all(s._cx_dev["pflam"] == 123.0))
    assert (int(torch.count_nonzero(la)) > 0) == bool(on)


def test_solver_warm_start_compact_plan_equals_int4_plan():
    """A learned-start plan (network every substep, the judge on every
    other frame) with compact rows equals the same plan without them."""
    import torch

    from xpbd_amd import warmstart as ws
    from xpbd_amd import Solver
    from xpbd_amd.scenes import cloth_sheet

    th = ws.init_params(5)
    th[-1] = -0.5
    theta = torch.from_numpy(th).to(DEV)

    def build():
        s, _ = cloth_sheet(Solver, res=48, solve_mode="jacobi", device=DEV, iterations=1)
        s.tile_size = 64
        s.tile_pair = False    # the start launch and the iteration as separate launches
        s.warm_start_judge = 2
        s.set_warm_start(theta)
        return s

    a, b = _pair(build, 40)
    assert a._cx_struct is not None and b._cx_struct is None and a._plan_pf
    _equal(a, b)
