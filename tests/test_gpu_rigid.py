"""GPU tests of the 6-DOF solver: parity of the HIP path with the CPU oracle
(bit-exact candidate pair sets and per-pair contact counts; body state per
substep from identical inputs within 1e-4, since the Jacobi accumulation
order differs), the broad phases against brute force, and the reference's
invariant tests (tests/test_solver_6dof.py) on the device."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
G = (0.0, -9.81, 0.0)


def _rel(g, o):
    g = np.asarray(g, np.float64)
    o = np.asarray(o, np.float64)
    return float(np.abs(g - o).max()) / max(float(np.abs(o).max()), 1e-30)


def _oracle6(oracle, s, broadphase=None):
    return oracle.Scene6(s.host_arrays(), dt=s.dt, substeps=s.substeps, iterations=s.iterations,
                         gravity=s.gravity, floor_y=s.floor_y, friction=s.friction,
                         restitution=s.restitution, lin_damp=s.lin_damp, ang_damp=s.ang_damp,
                         gyroscopic=s.gyroscopic, relax=s.jacobi_relax, max_speed=s.max_speed,
                         max_omega=s.max_omega, broadphase=broadphase or s.broadphase)


def _load_state(s, orc):
    import torch

    b = s._buf
    b["X"][:, :3] = torch.from_numpy(orc.x).to(DEV)
    b["Q"][:] = torch.from_numpy(orc.q).to(DEV)
    b["V"][:, :3] = torch.from_numpy(orc.v).to(DEV)
    b["W"][:, :3] = torch.from_numpy(orc.w).to(DEV)


def _random_boxes(n, seed, spread=3.0):
    from xpbd_amd import Solver6DOF
    from xpbd_amd.scenes import _random_quats

    rng = np.random.default_rng(seed)
    s = Solver6DOF(dt=1 / 60, substeps=1, iterations=1, device=DEV)
    q = _random_quats(rng, n)
    for i in range(n):
        he = rng.uniform(0.05, 0.2, 3)
        s.add_box(tuple(rng.uniform(-spread, spread, 3) + [0, spread + 1, 0]), tuple(he),
                  mass=0.0 if i % 17 == 0 else 1.0, quaternion=tuple(q[i]),
                  velocity=tuple(rng.normal(size=3)), omega=tuple(rng.normal(size=3)))
    return s


@pytest.mark.parametrize("broadphase", ["lbvh", "hashgrid"])
@pytest.mark.parametrize("n,seed", [(2, 0), (300, 1), (3000, 2), (6000, 3)])
def test_broadphase_pair_set_bit_exact(oracle, broadphase, n, seed):
    """The device pair list equals the oracle's brute-force list: same pairs,
    same (i, j) order, same count (LBVH: the all-pairs sweep up to 4096
    bodies, the tree above)."""
    s = _random_boxes(n, seed, spread=3.0 if n > 100 else 0.2)
    s.broadphase = broadphase
    s._flush()
    orc = _oracle6(oracle, s)
    orc.broadphase_pairs()
    cnt = s._run_broadphase(n, s.cap)
    if cnt > s.cap:
        s.cap = int(cnt * 1.5)
        s._alloc_pair_buffers(s.cap)
        cnt = s._run_broadphase(n, s.cap)
    s.n_pairs = cnt
    want = np.stack([orc.pair_a[:orc.n_pairs], orc.pair_b[:orc.n_pairs]], axis=1)
    assert cnt == orc.n_pairs
    assert (s.pairs() == want).all()


@pytest.mark.parametrize("n,seed", [(1, 0), (7, 4), (300, 1), (3000, 2)])
def test_lbvh_leaf_sizes_give_the_oracle_pairs(oracle, n, seed):
    """Every LBVH leaf size (1 .. 64 bodies per leaf, leaves larger than the
    scene included, and 0 = all pairs) emits the oracle's brute-force pair
    list: same pairs, same (i, j) order, same count."""
    s = _random_boxes(n, seed, spread=3.0 if n > 100 else 0.2)
    s._flush()
    orc = _oracle6(oracle, s)
    orc.broadphase_pairs()
    want = np.stack([orc.pair_a[:orc.n_pairs], orc.pair_b[:orc.n_pairs]], axis=1)
    for leaf in (0, 1, 2, 3, 8, 16, 64):
        s.lbvh_leaf = leaf
        cnt = s._run_broadphase(n, s.cap)
        if cnt > s.cap:
            s.cap = int(cnt * 1.5)
            s._alloc_pair_buffers(s.cap)
            cnt = s._run_broadphase(n, s.cap)
        s.n_pairs = cnt
        assert cnt == orc.n_pairs, leaf
        assert (s.pairs() == want).all(), leaf


def test_autotune_leaf_size_keeps_state_and_pairs():
    """The leaf-size autotuner times frames from one state, restores it, and
    the picked leaf size steps to the same pair list as the default."""
    from xpbd_amd import Solver6DOF, autotune
    from xpbd_amd.scenes import box_pile

    s = box_pile(Solver6DOF, 8, 5, 8, substeps=4, device=DEV)
    r = box_pile(Solver6DOF, 8, 5, 8, substeps=4, device=DEV)
    s.block_frames = r.block_frames = False
    s.step()
    r.step()
    x0 = s.positions()
    best, times = autotune.tune6(s, frames=2)
    assert best in times and 0 in times and 1 in times
    assert np.array_equal(s.positions(), x0)
    s.step()
    r.step()
    assert s.n_pairs == r.n_pairs and (s.pairs() == r.pairs()).all()
    with pytest.raises(RuntimeError):
        s.lbvh_leaf = 65
        s.step()


def test_broadphase_overflow_regrows():
    """A dense pile with a tiny pair capacity: the frame detects the overflow
    from the device count, grows the buffers and re-emits (solver_6dof.py:443-453)."""
    from xpbd_amd import Solver6DOF
    from xpbd_amd.scenes import box_pile

    s = box_pile(Solver6DOF, 6, 4, 6, substeps=4, device=DEV)
    s.block_frames = False          # the per-launch path sizes the pair list by the count
    s._flush()
    s.cap = 8
    s._alloc_pair_buffers(8)
    s.step()
    assert s.n_pairs > 8 and s.cap >= s.n_pairs
    assert len(s.pairs()) == s.n_pairs


@pytest.mark.parametrize("broadphase", ["lbvh", "hashgrid"])
def test_two_launch_plan_graph_equals_eager_launches(broadphase):
    """The frame plan with its broad phase inside, replayed from the
    captured graph and issued eagerly: the same pair lists and contact
    counts frame after frame, states within 1e-5 (float atomics sum in a
    different order)."""
    from xpbd_amd import Solver6DOF
    from xpbd_amd.scenes import box_pile

    a = box_pile(Solver6DOF, 5, 4, 5, device=DEV, broadphase=broadphase)
    b = box_pile(Solver6DOF, 5, 4, 5, device=DEV, broadphase=broadphase)
    for s, g in ((a, True), (b, False)):
        s.block_frames = False
        s.use_graph = g
    seen = 0
    for f in range(15):  # the pile's boxes start 4 cm apart: pairs come as it lands
        a.step()
        b.step()
        assert a.n_pairs == b.n_pairs, f
        assert (a.pairs() == b.pairs()).all(), f
        assert (a.contact_counts() == b.contact_counts()).all(), f
        assert _rel(a.positions(), b.positions()) <= 1e-5, f
        seen = max(seen, a.n_pairs)
    assert seen > 0


@pytest.mark.parametrize("use_graph", [True, False])
def test_overflowed_frames_run_again_when_read(use_graph):
    """The two-launch plan runs its broad phase inside the frame and checks
    the pair count later: a frame that overflowed the pair buffers, and
    every frame after it, wrote nothing; the next read of the state grows
    the buffers and runs those frames again.  Three frames stepped with a
    tiny capacity, then read, equal three frames with room from the start
    (pair lists exact, state within 1e-5: atomic sums in another order); a
    knob changed between frames settles the frames run before it with the
    old value."""
    from xpbd_amd import Solver6DOF
    from xpbd_amd.scenes import box_pile

    a = box_pile(Solver6DOF, 6, 4, 6, substeps=4, device=DEV)
    b = box_pile(Solver6DOF, 6, 4, 6, substeps=4, device=DEV)
    for s in (a, b):
        s.block_frames = False
        s.use_graph = use_graph
        s._flush()
    a.cap = 8
    a._alloc_pair_buffers(8)
    for _ in range(3):
        a.step()
        b.step()
    assert a.cap >= a.n_pairs == b.n_pairs > 8
    assert (a.pairs() == b.pairs()).all()
    assert _rel(a.positions(), b.positions()) <= 1e-5
    # overflow again, then a knob change before any read
    a.cap = 8
    a._alloc_pair_buffers(8)
    a.step()
    b.step()
    a.friction = b.friction = 0.3
    a.step()
    b.step()
    assert a.n_pairs == b.n_pairs
    assert _rel(a.positions(), b.positions()) <= 1e-5
    assert _rel(a.velocities(), b.velocities()) <= 1e-3


def test_morton_codes_decode(oracle):
    """Morton codes decode to the quantised coordinates (10 bits per axis,
    x in the highest bit of each triple)."""
    from xpbd_amd import _native as N

    lib = N.lib()
    rng = np.random.default_rng(0)
    for _ in range(200):
        p = rng.random(3).astype(np.float32)
        code = lib.xpbd_morton3(*map(float, p))
        q = np.minimum((p * 1024).astype(np.int64), 1023)
        dec = [0, 0, 0]
        for bit in range(10):
            for ax in range(3):
                dec[ax] |= ((code >> (3 * bit + 2 - ax)) & 1) << bit
        assert dec == q.tolist()
    assert lib.xpbd_morton3(0.0, 0.0, 0.0) == 0
    assert lib.xpbd_morton3(1.0, 1.0, 1.0) == (1 << 30) - 1


def _pile_scene(seed=0, gyroscopic=False):
    """60 random boxes in three layers (mu 0.5, e 0.2) plus a 5-link joint
    chain; with ``gyroscopic`` the boxes start spinning at up to 4 rad/s so
    the gyroscopic term of integrate_bodies is large."""
    from xpbd_amd import Solver6DOF
    from xpbd_amd.scenes import _random_quats

    rng = np.random.default_rng(seed)
    s = Solver6DOF(dt=1 / 60, substeps=8, iterations=2, device=DEV, friction=0.5,
                   restitution=0.2, gyroscopic=gyroscopic)
    q = _random_quats(rng, 60)
    k = 0
    for iy in range(3):
        for ix in range(5):
            for iz in range(4):
                he = rng.uniform(0.1, 0.25, 3)
                w = tuple(rng.uniform(-4, 4, 3)) if gyroscopic else (0.0, 0.0, 0.0)
                s.add_box((ix * 0.55, 0.3 + iy * 0.5, iz * 0.55), tuple(he),
                          mass=float(8000 * he.prod()), quaternion=tuple(q[k]), omega=w)
                k += 1
    a = s.add_box((4.0, 3.0, 0.0), (0.05, 0.05, 0.05), mass=0.0, static=True)
    prev = a
    for i in range(5):
        b = s.add_box((4.0 + 0.2 * (i + 1), 3.0, 0.0), (0.1, 0.03, 0.03), mass=0.5)
        s.add_joint(prev, b, compliance=1e-5)
        prev = b
    return s


def _set_mode(s, mode):
    """Frame path: the two-launch substep plan, the one-launch-per-phase
    plan, or the whole-frame workgroup kernel."""
    s.block_frames = mode == "block"
    s.substep_launches = 5 if mode == "five" else 2


MODES = ["two", "five", "block"]


@pytest.mark.parametrize("mode,gyro", [(m, g) for g in (False, True) for m in MODES])
def test_rigid_substep_parity_and_contact_counts(oracle, mode, gyro):
    """From identical inputs (the oracle's trajectory), one substep on the
    device matches the oracle: pair lists and per-pair contact counts
    bit-exact, x/q/v/w within 1e-4; through both per-launch plans and
    through the whole-frame workgroup kernel, with and without the
    gyroscopic term (kernels_6dof.py:91-101)."""
    s_full = _pile_scene(gyroscopic=gyro)
    orc = _oracle6(oracle, s_full)
    h = float(np.float32(1 / 60 / 8))
    s = _pile_scene(gyroscopic=gyro)
    _set_mode(s, mode)
    s.dt, s.substeps = h, 1
    s._flush()
    one = _oracle6(oracle, s)
    total_points = 0
    for k in range(6):
        orc.step(5)
        for name in ("x", "q", "v", "w"):
            getattr(one, name)[:] = getattr(orc, name)
        _load_state(s, one)
        s.step()
        one.step(1)
        assert s.n_pairs == one.n_pairs
        want = np.stack([one.pair_a[:one.n_pairs], one.pair_b[:one.n_pairs]], axis=1)
        assert (s.pairs() == want).all()
        mc = s.contact_counts()
        assert (mc == one.m_count[:one.n_pairs]).all(), k
        total_points += int(mc.sum())
        errs = [_rel(s.positions(), one.x), _rel(s.orientations(), one.q),
                _rel(s.velocities(), one.v), _rel(s.omega.numpy(), one.w)]
        assert max(errs) <= 1e-4, (k, errs)
    assert total_points > 0


@pytest.mark.parametrize("mode", MODES)
def test_rigid_frames_track_oracle(oracle, mode):
    """Whole frames (8 substeps x 2 iterations; graph replay, or one
    workgroup launch) over 20 frames of a 60-box pile.  The pile is chaotic:
    a one-ulp change of the oracle's own start positions grows to a 2.7e-3
    relative difference by frame 3 and to 0.18 by frame 20.  Each frame the
    device trajectory must stay within 1e-3, or within 4x that
    self-sensitivity of the oracle where it is larger."""
    s = _pile_scene(1)
    _set_mode(s, mode)
    h = s.host_arrays()
    orc = _oracle6(oracle, s)
    pert = _oracle6(oracle, s)
    pert.x = (h["x"] * np.float32(1 + 1.2e-7)).astype(np.float32)
    for f in range(20):
        s.step()
        orc.step(1)
        pert.step(1)
        e, sens = _rel(s.positions(), orc.x), _rel(pert.x, orc.x)
        assert e <= max(1e-3, 4 * sens), (f, e, sens)


def test_block_frames_match_launch_plan():
    """The whole-frame workgroup kernel against the per-launch plan on the
    100-box pile of the reference benchmark: identical pair lists and contact
    counts after the first frame, states within 1e-5 (float atomics sum in a
    different order), and a settled pile after 60 frames both ways."""
    from xpbd_amd import Solver6DOF
    from xpbd_amd.scenes import box_pile

    a = box_pile(Solver6DOF, 5, 4, 5, device=DEV, broadphase="lbvh")
    b = box_pile(Solver6DOF, 5, 4, 5, device=DEV, broadphase="lbvh")
    a.block_frames, b.block_frames = True, False
    a.step()
    b.step()
    assert a.n_pairs == b.n_pairs > 0
    assert (a.pairs() == b.pairs()).all()
    assert (a.contact_counts() == b.contact_counts()).all()
    assert _rel(a.positions(), b.positions()) <= 1e-5
    assert _rel(a.velocities(), b.velocities()) <= 1e-3
    for _ in range(59):
        a.step()
        b.step()
    for s in (a, b):
        x, v = s.positions(), s.velocities()
        assert np.all(np.isfinite(x))
        assert x[:, 1].min() > 0.16 - 0.02
        assert np.abs(v).max() < 0.5


@pytest.mark.parametrize("dims,iters", [((8, 6, 8), 1), ((5, 4, 5), 2)])
def test_two_launch_substep_matches_five_launch_plan(dims, iters):
    """The two-launch substep (per-body phases recomputed on values inside
    the contact and velocity launches) against one launch per phase on a
    reference benchmark pile: identical pair lists and contact counts after
    the first frame, states within 1e-5 after it (float atomics sum in a
    different order), both piles settled after 40 frames; one and two
    Jacobi iterations (the latter runs the extra apply + solve launches on
    the predicted pose)."""
    from xpbd_amd import Solver6DOF
    from xpbd_amd.scenes import box_pile

    a = box_pile(Solver6DOF, *dims, device=DEV, broadphase="lbvh", iterations=iters)
    b = box_pile(Solver6DOF, *dims, device=DEV, broadphase="lbvh", iterations=iters)
    for s, m in ((a, "two"), (b, "five")):
        _set_mode(s, m)
        s.step()
    assert a.n_pairs == b.n_pairs > 0
    assert (a.pairs() == b.pairs()).all()
    assert (a.contact_counts() == b.contact_counts()).all()
    assert _rel(a.positions(), b.positions()) <= 1e-5
    assert _rel(a.orientations(), b.orientations()) <= 1e-5
    for _ in range(39):
        a.step()
        b.step()
    for s in (a, b):
        x, v = s.positions(), s.velocities()
        assert np.all(np.isfinite(x))
        assert x[:, 1].min() > 0.16 - 0.02
        assert np.abs(v).max() < 1.0


# ---- the reference's invariant tests (tests/test_solver_6dof.py) -----------
def test_free_fall():
    from xpbd_amd import Solver6DOF

    s = Solver6DOF(dt=1 / 60, substeps=10, iterations=1, device=DEV, floor_y=-100.0)
    s.add_box((0, 0, 0), (0.2, 0.2, 0.2), mass=1.0)
    for _ in range(60):
        s.step()
    assert abs(s.positions()[0, 1] - 0.5 * -9.81) < 0.05
    assert abs(s.velocities()[0, 1] - (-9.81)) < 0.05


def test_box_rests_on_floor():
    from xpbd_amd import Solver6DOF

    s = Solver6DOF(dt=1 / 60, substeps=15, iterations=1, device=DEV, floor_y=0.0, friction=0.6)
    s.add_box((0, 1.0, 0), (0.2, 0.2, 0.2), mass=1.0)
    for _ in range(180):
        s.step()
    assert abs(s.positions()[0, 1] - 0.2) < 2e-3
    assert abs(s.velocities()[0, 1]) < 1e-2
    assert abs(abs(s.orientations()[0, 3]) - 1.0) < 1e-2


def test_spin_conservation():
    from xpbd_amd import Solver6DOF

    s = Solver6DOF(dt=1 / 60, substeps=10, iterations=1, device=DEV, gravity=(0, 0, 0),
                   floor_y=-100.0)
    s.add_box((0, 0, 0), (0.2, 0.1, 0.3), mass=1.0)
    s._flush()
    om = s.omega.numpy().copy()
    om[0] = [0, 5.0, 0]
    s.omega = om
    s._graph = None
    for _ in range(120):
        s.step()
    w = s.omega.numpy()[0]
    assert abs(w[1] - 5.0) < 0.5
    assert np.linalg.norm(w[[0, 2]]) < 0.5


@pytest.mark.parametrize("n", [3, 5])
def test_stack_settles_no_penetration(n):
    from xpbd_amd import Solver6DOF

    s = Solver6DOF(dt=1 / 60, substeps=20, iterations=1, device=DEV, floor_y=0.0, friction=0.6)
    h = 0.2
    for k in range(n):
        s.add_box((0, h + k * (2 * h + 0.02), 0), (h, h, h), mass=1.0)
    for _ in range(350):
        s.step()
    ys = np.sort(s.positions()[:, 1])
    assert np.all(np.isfinite(s.positions()))
    assert ys.max() < n * (2 * h) + 0.2
    assert np.allclose(np.diff(ys), 2 * h, atol=8e-3)
    assert np.abs(s.velocities()).max() < 0.1


def test_pile_stable():
    from xpbd_amd import Solver6DOF

    s = Solver6DOF(dt=1 / 60, substeps=18, iterations=1, device=DEV, floor_y=0.0, friction=0.6)
    h = 0.16
    rng = np.random.default_rng(0)
    for ly in range(3):
        for ix in range(3):
            for iz in range(3):
                s.add_box((ix * 0.34 + rng.uniform(-2e-3, 2e-3), h + ly * 0.34,
                           iz * 0.34 + rng.uniform(-2e-3, 2e-3)), (h, h, h), mass=1.0)
    for _ in range(400):
        s.step()
    p = s.positions()
    assert np.all(np.isfinite(p))
    assert p[:, 1].max() < 1.2
    assert np.abs(s.velocities()).max() < 0.2


def test_box_tips_and_rests_flat():
    from xpbd_amd import Solver6DOF

    s = Solver6DOF(dt=1 / 60, substeps=20, iterations=1, device=DEV, floor_y=0.0, friction=0.7)
    hw, hh, hd = 0.05, 0.25, 0.12
    s.add_box((0, hh, 0), (hw, hh, hd), mass=1.0)
    s._flush()
    om = s.omega.numpy().copy()
    om[0] = [0, 0, -3.0]
    s.omega = om
    s._graph = None
    for _ in range(300):
        s.step()
    assert abs(s.positions()[0, 1] - hw) < 0.03
    assert np.linalg.norm(s.omega.numpy()[0]) < 0.2


def test_broadphases_agree():
    from xpbd_amd import Solver6DOF

    def run(bp):
        s = Solver6DOF(dt=1 / 60, substeps=20, iterations=1, device=DEV, floor_y=0.0,
                       friction=0.6, broadphase=bp)
        h = 0.2
        for k in range(4):
            s.add_box((0, h + k * (2 * h + 0.02), 0), (h, h, h), mass=1.0)
        for _ in range(300):
            s.step()
        return np.sort(s.positions()[:, 1])

    ref = run("lbvh")
    assert np.allclose(run("grid"), ref, atol=5e-3)
    assert np.allclose(run("hashgrid"), ref, atol=5e-3)


def test_joint_rigid_link():
    from xpbd_amd import Solver6DOF

    s = Solver6DOF(dt=1 / 60, substeps=20, iterations=2, device=DEV, floor_y=-100.0)
    anchor = s.add_box((0, 2, 0), (0.05, 0.05, 0.05), mass=0.0, static=True)
    bob = s.add_box((0.6, 2, 0), (0.1, 0.1, 0.1), mass=1.0)
    s.add_joint(anchor, bob, compliance=0.0)
    for _ in range(300):
        s.step()
    p = s.positions()
    assert np.all(np.isfinite(p))
    assert abs(np.linalg.norm(p[bob.index] - p[anchor.index]) - 0.6) < 0.02
    assert p[bob.index, 1] < p[anchor.index, 1]


def test_particle_rests_on_floor():
    from xpbd_amd import Solver6DOF

    s = Solver6DOF(dt=1 / 60, substeps=15, iterations=1, device=DEV, floor_y=0.0, friction=0.5)
    r = 0.05
    s.add_particle((0, 1.0, 0), mass=0.1, radius=r)
    for _ in range(200):
        s.step()
    assert abs(s.positions()[0, 1] - r) < 5e-3
    assert abs(s.velocities()[0, 1]) < 1e-2


def test_cloth_hangs_from_pins():
    from xpbd_amd import Solver6DOF

    s = Solver6DOF(dt=1 / 60, substeps=15, iterations=2, device=DEV, floor_y=-100.0)
    R, sp = 8, 0.1
    idx = np.empty((R, R), np.int64)
    for iy in range(R):
        for ix in range(R):
            pinned = iy == 0
            b = s.add_particle((ix * sp, 2.0, iy * sp), mass=0.0 if pinned else 0.02,
                               radius=0.01, group=1, static=pinned)
            idx[iy, ix] = b.index
    for iy in range(R):
        for ix in range(R):
            if ix + 1 < R:
                s.add_joint(int(idx[iy, ix]), int(idx[iy, ix + 1]), compliance=1e-7)
            if iy + 1 < R:
                s.add_joint(int(idx[iy, ix]), int(idx[iy + 1, ix]), compliance=1e-7)
    for _ in range(250):
        s.step()
    P = s.positions()
    assert np.all(np.isfinite(P))
    assert abs(P[idx[0]][:, 1].mean() - 2.0) < 1e-3
    assert P[idx[R - 1]][:, 1].mean() < 1.85
    assert abs(np.linalg.norm(P[idx[4, 1]] - P[idx[4, 0]]) - sp) < 0.02


def test_unified_rigid_cloth_joints():
    from xpbd_amd import Solver6DOF

    s = Solver6DOF(dt=1 / 60, substeps=18, iterations=2, device=DEV, floor_y=0.0, friction=0.6,
                   lin_damp=0.01)
    for k in range(3):
        s.add_box((0, 0.2 + k * 0.42, 0), (0.2, 0.2, 0.2), mass=1.0)
    R, sp = 6, 0.12
    idx = np.empty((R, R), np.int64)
    for iy in range(R):
        for ix in range(R):
            pinned = iy == 0
            b = s.add_particle((1.5 + ix * sp, 1.5, iy * sp), mass=0.0 if pinned else 0.02,
                               radius=0.01, group=1, static=pinned)
            idx[iy, ix] = b.index
    for iy in range(R):
        for ix in range(R):
            if ix + 1 < R:
                s.add_joint(int(idx[iy, ix]), int(idx[iy, ix + 1]), compliance=1e-6)
            if iy + 1 < R:
                s.add_joint(int(idx[iy, ix]), int(idx[iy + 1, ix]), compliance=1e-6)
    assert s.num_joints > 0
    for _ in range(300):
        s.step()
    P = s.positions()
    assert np.all(np.isfinite(P))
    ys = np.sort(P[:3, 1])
    assert np.allclose(np.diff(ys), 0.4, atol=1e-2)
    assert np.abs(s.velocities()).max() < 1.0


def test_domino_cascade():
    from xpbd_amd import Solver6DOF

    s = Solver6DOF(dt=1 / 60, substeps=20, iterations=1, device=DEV, floor_y=0.0, friction=0.6)
    hw, hh, hd = 0.04, 0.22, 0.12
    sp = 0.18
    n = 6
    ang = -0.45
    q0 = (0.0, 0.0, float(np.sin(ang / 2)), float(np.cos(ang / 2)))
    for k in range(n):
        q = q0 if k == 0 else (0, 0, 0, 1)
        y = hh * 0.9 if k == 0 else hh
        s.add_box((k * sp, y, 0.0), (hw, hh, hd), mass=1.0, quaternion=q)
    for _ in range(420):
        s.step()
    fell = int(np.sum(s.positions()[:, 1] < hh * 0.6))
    assert fell >= n - 1


def test_unified_config3_runs_and_settles_contacts(oracle):
    """Config 3 (1440 boxes + 32^2 cloth + 20-link chain, LBVH, 15x2): the
    first frame's pair list equals the oracle's; frames run through the
    captured graph and stay finite; then, once the boxes have landed (frames
    40, 70 and 100), one substep from the device's state on a one-substep
    copy of the scene and in the oracle from the same state: pair lists and
    per-pair manifold contact counts bit-exact, x/q/v/w within 1e-4."""
    from xpbd_amd import Solver6DOF
    from xpbd_amd.scenes import unified_scene

    s = unified_scene(Solver6DOF, device=DEV)
    orc = _oracle6(oracle, s)
    s.step()
    orc.step(1)
    want = np.stack([orc.pair_a[:orc.n_pairs], orc.pair_b[:orc.n_pairs]], axis=1)
    assert s.n_pairs == orc.n_pairs and (s.pairs() == want).all()
    h = float(np.float32(s.dt / s.substeps))
    one = unified_scene(Solver6DOF, device=DEV, dt=h, substeps=1)
    one._flush()
    ref = _oracle6(oracle, one)
    done, points = 1, 0
    for target in (40, 70, 100):
        while done < target:
            s.step()
            done += 1
        assert np.all(np.isfinite(s.positions()))
        st = {k: s._buf[k].clone() for k in ("X", "Q", "V", "W")}
        for k in ("X", "Q", "V", "W"):
            one._buf[k].copy_(st[k])
        ref.x[:] = st["X"][:, :3].cpu().numpy()
        ref.q[:] = st["Q"].cpu().numpy()
        ref.v[:] = st["V"][:, :3].cpu().numpy()
        ref.w[:] = st["W"][:, :3].cpu().numpy()
        one.step()
        ref.step(1)
        assert one.n_pairs == ref.n_pairs, target
        want = np.stack([ref.pair_a[:ref.n_pairs], ref.pair_b[:ref.n_pairs]], axis=1)
        assert (one.pairs() == want).all(), target
        mc = one.contact_counts()
        assert (mc == ref.m_count[:ref.n_pairs]).all(), target
        points += int(mc.sum())
        errs = [_rel(one.positions(), ref.x), _rel(one.orientations(), ref.q),
                _rel(one.velocities(), ref.v), _rel(one.omega.numpy(), ref.w)]
        assert max(errs) <= 1e-4, (target, errs)
    assert points > 500
    print(f"config3: {s.num_bodies} bodies, {s.num_joints} joints, {s.n_pairs} pairs, "
          f"{points} manifold points over the three checks")
