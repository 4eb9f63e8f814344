"""GPU parity of the 3-DOF CONTACT rows (sphere-sphere self-collision,
src/xpbd3d/kernels.py:182-208 for Gauss-Seidel, 294-322 for Jacobi) against
the CPU oracle, on the reference benchmark's sphere stack
(examples/benchmark.py:93-107) and on a dense sphere pile:

* host rows (the reference's per-frame broad phase, solver.py:275-342), GS:
  frame by frame over a trajectory, every state bit-exact (tolerance written
  1e-5), through the per-color launches and the whole-frame workgroup;
* host rows, Jacobi: per substep from identical inputs within 1e-4;
* device rows (``xpbd_sphere_contacts`` projected by
  ``k_jacobi_project_dyn``): per frame of one or two substeps from identical
  inputs within 1e-4, including the projection from predicted positions
  (integrate folded into the first iteration) and the plain projection after
  the previous substep's folded integrate; the rows the oracle gets are the
  rows the device generated, read back and mapped to API indices.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
CONTACT = 3


def _rel(g, o):
    g = np.asarray(g, np.float64)
    o = np.asarray(o, np.float64)
    return float(np.abs(g - o).max()) / max(float(np.abs(o).max()), 1e-30)


def _oracle_scene(oracle, s, h, x, v, dt=None, substeps=None):
    hh = dict(h)
    hh["x"], hh["v"] = np.asarray(x, np.float32), np.asarray(v, np.float32)
    return oracle.Scene3(hh, dt=s.dt if dt is None else dt,
                         substeps=s.substeps if substeps is None else substeps,
                         iterations=s.iterations, gravity=s.gravity, mode=s.solve_mode,
                         damping=s.damping, relax=s.jacobi_relax, max_speed=s.max_speed)


def _stack(mode, **kw):
    from xpbd_amd import Solver
    from xpbd_amd.scenes import sphere_stack

    return sphere_stack(Solver, 8, 5, mode, device=DEV, **kw)


def _pile(mode="jacobi", k=343, seed=5, moving=False, **kw):
    """A jittered cube of k spheres (r = 0.05 m) 2 mm apart on the floor,
    mu 0.5; ``moving`` adds random velocities (a larger search margin)."""
    from xpbd_amd import Solver

    s = Solver(dt=1 / 60, substeps=10, iterations=2, solve_mode=mode, device=DEV, **kw)
    rng = np.random.default_rng(seed)
    side = int(round(k ** (1 / 3)))
    g = np.arange(k)
    ix, rem = g % side, g // side
    iz, iy = rem % side, rem // side
    pos = np.stack([ix * 0.102, 0.06 + iy * 0.102, iz * 0.102], 1)
    pos += rng.uniform(-3e-3, 3e-3, pos.shape)
    vel = rng.normal(0, 0.4, pos.shape) if moving else None
    s.add_particles(pos, np.ones(k), velocities=vel, radius=0.05, collide=True, friction=0.5)
    s.add_floor_contacts(np.arange(k), floor_y=0.05)
    return s


@pytest.mark.parametrize("block", [False, True])
def test_stack_gs_host_contact_rows_bit_exact_per_frame(oracle, block):
    """The 8 x 8 x 5 sphere stack in GS (15 x 2), 45 frames: each frame the
    solver rebuilds its CONTACT rows on the host, the oracle steps the same
    frame from its own state with those rows (recolored by the oracle's own
    greedy coloring); positions, velocities and multipliers must agree to
    1e-5 every frame, and are expected bit for bit."""
    s = _stack("gs")
    s.block_frames = block
    x = s.positions().copy()
    v = s.velocities().copy()
    exact = True
    n_contact_frames = 0
    for f in range(45):
        s.step()
        h = s.host_arrays()
        n_ct = int((h["type"] == CONTACT).sum())
        n_contact_frames += n_ct > 0
        orc = _oracle_scene(oracle, s, h, x, v)
        orc.step(1)
        x, v = orc.x.copy(), orc.v.copy()
        xg, vg, lg = s.positions(), s.velocities(), s.lambdas()
        errs = (_rel(xg, orc.x), _rel(vg, orc.v), _rel(lg, orc.lam))
        assert max(errs) <= 1e-5, (f, n_ct, errs)
        exact &= bool((xg == orc.x).all() and (vg == orc.v).all() and (lg == orc.lam).all())
    assert s._block_active == block
    assert n_contact_frames >= 30
    print(f"stack gs block={block}: bit-exact={exact}, {n_contact_frames} frames with contacts")
    assert exact


def _warm_states(build_full, frames_between, n_checks):
    s = build_full()
    out = []
    for _ in range(n_checks):
        s.step(frames_between)
        out.append((s.positions().copy(), s.velocities().copy()))
    return out


@pytest.mark.parametrize("scene,fuse", [("stack", True), ("stack", False), ("pile", True)])
def test_jacobi_host_contact_rows_per_substep(oracle, scene, fuse):
    """Jacobi with host-built CONTACT rows: from states of the full 15 x 2
    (pile: 10 x 2) trajectory, one substep on the device and in the oracle
    from identical inputs and the same rows agree within 1e-4 (positions,
    velocities, every row's multiplier)."""
    def build(one):
        s = _stack("jacobi") if scene == "stack" else _pile("jacobi")
        if scene == "pile":
            s.enable_self_collision(True, default_friction=0.5)
        if one:
            s.dt = float(np.float32(s.dt / s.substeps))
            s.substeps = 1
        s.fuse_first = fuse
        return s

    states = _warm_states(lambda: build(False), 4, 5)
    one = build(True)
    worst, total = 0.0, 0
    for k, (x, v) in enumerate(states):
        one.set_state(x, v)
        one.step()
        h = one.host_arrays()
        n_ct = int((h["type"] == CONTACT).sum())
        total += n_ct
        orc = _oracle_scene(oracle, one, h, x, v)
        orc.step(1)
        errs = (_rel(one.positions(), orc.x), _rel(one.velocities(), orc.v),
                _rel(one.lambdas(), orc.lam))
        worst = max(worst, max(errs))
        assert max(errs) <= 1e-4, (k, n_ct, errs)
    assert total > 200
    print(f"{scene} jacobi host rows fuse={fuse}: worst rel err {worst:.2e}, {total} contact rows")


def _device_rows(s):
    """This frame's device CONTACT rows in API indices, as oracle rows."""
    cnt = int(s.n_contacts)
    rows = s.CROWS[:cnt].cpu().numpy()
    a = (rows[:, 0] & 0x0FFFFFFF).astype(np.int64)
    b = rows[:, 1].astype(np.int64)
    assert np.all((rows[:, 0] >> 28) == CONTACT)
    if s._pperm is not None:
        a, b = s._pperm[a], s._pperm[b]
    return dict(a=a.astype(np.int32), b=b.astype(np.int32),
                rest=rows[:, 2].copy().view(np.float32), mu=rows[:, 3].copy().view(np.float32))


def _with_rows(h, r):
    m = len(r["a"])
    out = dict(h)
    out["type"] = np.r_[h["type"], np.full(m, CONTACT, np.int32)]
    out["a"] = np.r_[h["a"], r["a"]]
    out["b"] = np.r_[h["b"], r["b"]]
    out["rest"] = np.r_[h["rest"], r["rest"]]
    out["compliance"] = np.r_[h["compliance"], np.zeros(m, np.float32)]
    out["anchor"] = np.r_[h["anchor"].reshape(-1, 3), np.zeros((m, 3), np.float32)]
    out["mu"] = np.r_[h["mu"], r["mu"]]
    return out


@pytest.mark.parametrize("scene,substeps,fuse", [
    ("pile", 1, True),        # first iteration from predicted positions (dyn_first)
    ("pile", 2, True),        # + plain projection after the previous substep's fold
    ("pile", 2, False),       # separate integrate pass
    ("moving", 2, True),      # random velocities: wider search margin, friction sliding
    ("stack", 2, True),       # the reference benchmark's stack
])
def test_device_contact_rows_projection_matches_oracle(oracle, scene, substeps, fuse):
    """Device-generated CONTACT rows through k_jacobi_project_dyn: from
    states of the full trajectory, one frame of ``substeps`` substeps on the
    device and in the oracle (static rows + the device's rows, in the same
    order) from identical inputs: positions, velocities, the static rows'
    and the contact rows' multipliers within 1e-4."""
    def build(frame_substeps):
        if scene == "stack":
            s = _stack("jacobi")
            s.enable_self_collision(False)
        else:
            s = _pile("jacobi", moving=scene == "moving")
        h = float(np.float32(s.dt / s.substeps))
        if frame_substeps:
            s.dt = h * frame_substeps
            s.substeps = frame_substeps
        s.enable_device_contacts(True, margin=0.01)
        s.fuse_first = fuse
        return s

    states = _warm_states(lambda: build(None), 3, 5)
    one = build(substeps)
    worst, total = 0.0, 0
    for k, (x, v) in enumerate(states):
        one.set_state(x, v)
        one.step()
        r = _device_rows(one)
        total += len(r["a"])
        h = _with_rows(one.host_arrays(), r)
        orc = _oracle_scene(oracle, one, h, x, v)
        orc.step(1)
        n_static = one.num_constraints
        lam_dyn = one.CLAM[:len(r["a"])].cpu().numpy()
        errs = (_rel(one.positions(), orc.x), _rel(one.velocities(), orc.v),
                _rel(one.lambdas(), orc.lam[:n_static]),
                _rel(lam_dyn, orc.lam[n_static:]) if len(r["a"]) else 0.0)
        worst = max(worst, max(errs))
        assert max(errs) <= 1e-4, (k, len(r["a"]), errs)
    assert total > 300
    print(f"{scene} device rows substeps={substeps} fuse={fuse}: worst {worst:.2e}, "
          f"{total} contact rows")


def test_live_knobs_rewrite_rows_in_place_and_keep_the_graph(oracle):
    """set_all_compliance / set_all_friction on a captured tile-mode plan
    (a 32^2 cloth on the floor plus the device CONTACT generator of a sphere
    pile in the same solver): the row buffers and the friction the device
    contact rows are generated from are rewritten in place, the captured
    plan is kept (same handle), the new mu reaches the next frame's device
    rows, and the following frames match the oracle stepped with the new
    values per substep within 1e-4."""
    s = _pile("jacobi", k=125)
    s.enable_device_contacts(True, margin=0.01)
    s.block_frames = False
    s.step(3)
    plan = s._plan.value
    ptrs = (s.R.data_ptr(), s.CMU.data_ptr(), s.CROWS.data_ptr())
    s.set_all_friction(0.15)
    s.set_all_compliance(1e-7, only_distance=False)
    s.step()
    assert s._plan.value == plan
    assert (s.R.data_ptr(), s.CMU.data_ptr(), s.CROWS.data_ptr()) == ptrs
    assert np.all(s.CMU.cpu().numpy() == np.float32(0.15))
    r = _device_rows(s)
    assert len(r["mu"]) > 0 and np.all(r["mu"] == np.float32(0.15))
    h = s.host_arrays()
    assert np.all(h["mu"][h["type"] == 2] == np.float32(0.15))
    # one more frame against the oracle from the device state, same rows
    x, v = s.positions(), s.velocities()
    s.step()
    r = _device_rows(s)
    orc = _oracle_scene(oracle, s, _with_rows(h, r), x, v)
    orc.step(1)
    assert _rel(s.positions(), orc.x) <= 1e-4


# ---- device colouring and Gauss-Seidel with device CONTACT rows -------------
def _random_row_graph(seed, n, m, one_body_frac=0.2):
    rng = np.random.default_rng(seed)
    a = rng.integers(0, n, m)
    b = rng.integers(0, n, m)
    b = np.where(b == a, (b + 1) % n, b)
    b = np.where(rng.random(m) < one_body_frac, -1, b)
    return a.astype(np.int32), b.astype(np.int32)


@pytest.mark.parametrize("seed,n,m", [(0, 50, 40), (1, 300, 2000), (2, 4096, 12000),
                                      (3, 64, 1500)])
def test_device_coloring_equals_the_reference_greedy(oracle, seed, n, m):
    """xpbd_color_rows_device (Jones-Plassmann with the greedy order as
    priority, one workgroup) gives every row the colour the reference's
    sequential greedy colouring gives it (the oracle's restatement of
    coloring.py:26-71), split across static rows and a device-counted
    segment; the colour ranges hold every row once, rows of one colour share
    no body."""
    import torch

    from xpbd_amd import _native as N

    a, b = _random_row_graph(seed, n, m)
    want = oracle.color_constraints(n, a, b)
    R = np.zeros((m, 4), np.int32)
    R[:, 0], R[:, 1] = a, b
    m_static = m // 2
    dev = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to(DEV)  # noqa: E731
    Rd = dev(R)
    cnt = dev(np.array([m - m_static], np.int32))
    inc = torch.empty(2 * m, dtype=torch.int32, device=DEV)
    col = torch.empty(m, dtype=torch.int32, device=DEV)
    order = torch.empty(m, dtype=torch.int32, device=DEV)
    starts = torch.empty(257, dtype=torch.int32, device=DEV)
    nc = torch.empty(1, dtype=torch.int32, device=DEV)
    N.check(N.lib().xpbd_color_rows_device(None, N.ptr(Rd), m_static, N.ptr(cnt), m - m_static,
                                           n, N.ptr(inc), N.ptr(col), N.ptr(order),
                                           N.ptr(starts), N.ptr(nc)), "color")
    torch.cuda.synchronize()
    got = col.cpu().numpy()
    k = int(nc.item())
    assert np.array_equal(got, want)
    assert k == int(want.max()) + 1
    st, od = starts.cpu().numpy()[:k + 1], order.cpu().numpy()
    assert st[0] == 0 and st[-1] == m and np.array_equal(np.sort(od), np.arange(m))
    for c in range(k):
        rows = od[st[c]:st[c + 1]]
        assert np.all(got[rows] == c)
        bodies = np.r_[a[rows], b[rows][b[rows] >= 0]]
        assert len(np.unique(bodies)) == len(bodies)


@pytest.mark.parametrize("scene", ["stack", "pile"])
def test_gs_device_contacts_bit_exact_with_oracle_and_host_rows(oracle, scene):
    """Gauss-Seidel with device CONTACT rows (rows generated, recoloured and
    solved on the device, the frame in one workgroup) against (1) the oracle
    stepped frame by frame from its own state with the rows the device
    generated, and (2) the host self-collision path (rows rebuilt and
    recoloured on the host every frame): positions, velocities and
    multipliers bit for bit over 40 frames."""
    def build(device):
        if scene == "stack":
            s = _stack("gs")
            if device:
                s.enable_self_collision(False)
        else:
            s = _pile("gs", k=216)
            if not device:
                s.enable_self_collision(True, default_friction=0.5)
        if device:
            s.enable_device_contacts(True)
        return s

    d, hst = build(True), build(False)
    x, v = d.positions().copy(), d.velocities().copy()
    n_ct = 0
    for f in range(40):
        d.step()
        hst.step()
        r = d.device_contact_rows()
        n_ct += len(r["a"])
        h = d.host_arrays()
        m = len(r["a"])
        h = dict(h, type=np.r_[h["type"], np.full(m, CONTACT, np.int32)],
                 a=np.r_[h["a"], r["a"].astype(np.int32)],
                 b=np.r_[h["b"], r["b"].astype(np.int32)],
                 rest=np.r_[h["rest"], r["rest"]], mu=np.r_[h["mu"], r["mu"]],
                 compliance=np.r_[h["compliance"], np.zeros(m, np.float32)],
                 anchor=np.r_[h["anchor"].reshape(-1, 3), np.zeros((m, 3), np.float32)])
        orc = _oracle_scene(oracle, d, h, x, v)
        orc.step(1)
        x, v = orc.x.copy(), orc.v.copy()
        n_static = d.num_constraints
        lam_dev = np.r_[d.lambdas(), d.CLAM[:m].cpu().numpy()]
        assert np.array_equal(d.positions(), orc.x), f
        assert np.array_equal(d.velocities(), orc.v), f
        assert np.array_equal(lam_dev, orc.lam), f
        assert d.num_colors == orc.num_colors, f
        assert np.array_equal(d.positions(), hst.positions()), f
        assert np.array_equal(d.lambdas(), hst.lambdas()[:n_static]), f
    assert n_ct > 500
    print(f"{scene}: {n_ct} device contact rows over 40 frames, bit-exact")
