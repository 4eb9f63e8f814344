"""Pins the 6-DOF CPU oracle (oracle/oracle6.c) against textbook identities,
closed forms and the reference's own invariant tests
(tests/test_solver_6dof.py), run on the oracle."""
import numpy as np
import pytest

G = (0.0, -9.81, 0.0)


def _rot(q):
    """Textbook rotation matrix of a unit quaternion (x, y, z, w), float64."""
    x, y, z, w = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
        [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
        [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def _qaa(axis, ang):
    axis = np.asarray(axis, float) / np.linalg.norm(axis)
    s = np.sin(ang / 2)
    return np.array([axis[0] * s, axis[1] * s, axis[2] * s, np.cos(ang / 2)])


def _box_inv_inertia(he, m):
    lx, ly, lz = 2 * he[0], 2 * he[1], 2 * he[2]
    return (12 / (m * (ly * ly + lz * lz)), 12 / (m * (lx * lx + lz * lz)),
            12 / (m * (lx * lx + ly * ly)))


def _scene(bodies, joints=(), **kw):
    """bodies: list of dict(x, he, m, q=(0,0,0,1), v, w, particle, static, group)."""
    n = len(bodies)
    h = dict(x=np.zeros((n, 3)), q=np.zeros((n, 4)), v=np.zeros((n, 3)), w=np.zeros((n, 3)),
             inv_mass=np.zeros(n), inv_I=np.zeros((n, 3)), he=np.zeros((n, 3)),
             group=np.zeros(n, np.int32))
    for i, b in enumerate(bodies):
        h["x"][i] = b["x"]
        h["q"][i] = b.get("q", (0, 0, 0, 1))
        h["v"][i] = b.get("v", (0, 0, 0))
        h["w"][i] = b.get("w", (0, 0, 0))
        h["he"][i] = b["he"]
        h["group"][i] = b.get("group", 0)
        if b.get("static") or b["m"] <= 0:
            continue
        h["inv_mass"][i] = 1.0 / b["m"]
        if not b.get("particle"):
            h["inv_I"][i] = _box_inv_inertia(b["he"], b["m"])
    if joints:
        h["j_a"] = [j[0] for j in joints]
        h["j_b"] = [j[1] for j in joints]
        h["j_rest"] = [j[2] for j in joints]
        h["j_compliance"] = [j[3] for j in joints]
    return h


def test_aabb_half_extents_match_textbook_rotation(oracle):
    """compute_body_aabb: half = |R| he (+ margin pad), R from the textbook
    quaternion matrix; pins the restated quaternion builtins."""
    rng = np.random.default_rng(0)
    n = 32
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    he = rng.uniform(0.1, 0.3, size=(n, 3))
    x = rng.normal(size=(n, 3))
    f = lambda a: np.ascontiguousarray(a, np.float32)  # noqa: E731
    lo = np.zeros((n, 3), np.float32)
    hi = np.zeros((n, 3), np.float32)
    lb = oracle.lib6()
    X, Q, Z, HE = f(x), f(q), f(np.zeros((n, 3))), f(he)
    lb.orc6_body_aabb(X.ctypes.data, Q.ctypes.data, Z.ctypes.data, Z.ctypes.data,
                      HE.ctypes.data, 0.02, 1 / 60, lo.ctypes.data, hi.ctypes.data, n)
    for i in range(n):
        half = np.abs(_rot(Q[i].astype(float))) @ HE[i].astype(float) + 0.02
        np.testing.assert_allclose(hi[i] - X[i], half, rtol=2e-6, atol=2e-7)
        np.testing.assert_allclose(X[i] - lo[i], half, rtol=2e-6, atol=2e-7)


def test_free_fall_6dof(oracle):
    h = _scene([dict(x=(0, 0, 0), he=(0.2, 0.2, 0.2), m=1.0)])
    sc = oracle.Scene6(h, dt=1 / 60, substeps=10, iterations=1, gravity=G, floor_y=-100.0)
    sc.step(60)
    assert abs(sc.x[0, 1] - 0.5 * -9.81) < 0.05
    assert abs(sc.v[0, 1] - (-9.81)) < 0.05


def test_box_rests_on_floor(oracle):
    h = _scene([dict(x=(0, 1.0, 0), he=(0.2, 0.2, 0.2), m=1.0)])
    sc = oracle.Scene6(h, dt=1 / 60, substeps=15, iterations=1, gravity=G, friction=0.6)
    sc.step(180)
    assert abs(sc.x[0, 1] - 0.2) < 2e-3
    assert abs(sc.v[0, 1]) < 1e-2
    assert abs(abs(sc.q[0, 3]) - 1.0) < 1e-2


def test_spin_about_principal_axis_is_conserved(oracle):
    h = _scene([dict(x=(0, 0, 0), he=(0.2, 0.1, 0.3), m=1.0, w=(0, 5.0, 0))])
    sc = oracle.Scene6(h, dt=1 / 60, substeps=10, iterations=1, gravity=(0, 0, 0),
                       floor_y=-100.0)
    sc.step(120)
    assert abs(sc.w[0, 1] - 5.0) < 0.5
    assert np.linalg.norm(sc.w[0, [0, 2]]) < 0.5
    # the orientation after t seconds is a rotation of 5 t rad about y
    t = 2.0
    ang = np.arctan2(2 * sc.q[0, 1] * sc.q[0, 3], 1 - 2 * sc.q[0, 1] ** 2)
    assert np.cos(ang) == pytest.approx(np.cos(5 * t), abs=0.05)


def _manifold(oracle, bodies, margin=0.02):
    h = _scene(bodies)
    sc = oracle.Scene6(h, dt=1 / 60, substeps=1, iterations=1, gravity=G)
    sc.pair_a[0], sc.pair_b[0] = 0, 1
    sc.n_pairs = 1
    lb = oracle.lib6()
    P = oracle._P
    lb.orc6_box_manifold(P(sc.x), P(sc.q), P(sc.he), P(sc.pair_a), P(sc.pair_b), 1, margin,
                         P(sc.poly), P(sc.m_count), P(sc.m_inc), P(sc.m_ref), P(sc.m_normal),
                         P(sc.m_off_inc), P(sc.m_off_ref), sc.cap)
    return sc


def test_sat_manifold_aligned_stack_gives_four_face_corners(oracle):
    """An axis-aligned box resting 1 mm into an equal box below: the SAT
    normal is +y, the manifold is the 4 corners of the incident face."""
    sc = _manifold(oracle, [dict(x=(0, 0, 0), he=(0.2, 0.2, 0.2), m=1.0),
                            dict(x=(0, 0.399, 0), he=(0.2, 0.2, 0.2), m=1.0)])
    assert sc.m_count[0] == 4
    n = sc.m_normal[0]
    inc = sc.m_inc[0]
    # the pushed (incident) body moves along +n, away from the reference body
    other = 1 - inc
    assert np.dot(sc.x[inc] - sc.x[other], n) > 0
    assert abs(abs(n[1]) - 1.0) < 1e-6
    offs = np.sort(np.round(sc.m_off_inc[:4], 5), axis=0)
    assert np.allclose(np.abs(offs[:, [0, 2]]), 0.2, atol=1e-5)
    assert np.allclose(np.abs(offs[:, 1]), 0.2, atol=1e-5)


def test_sat_manifold_separated_is_empty(oracle):
    sc = _manifold(oracle, [dict(x=(0, 0, 0), he=(0.2, 0.2, 0.2), m=1.0),
                            dict(x=(0, 0.5, 0), he=(0.2, 0.2, 0.2), m=1.0)])
    assert sc.m_count[0] == 0


def test_sat_manifold_yawed_box_clips_to_octagon(oracle):
    """A box yawed 45 deg about the contact normal: the incident square
    clipped against the reference square is a regular octagon (8 points)."""
    q = _qaa((0, 1, 0), np.pi / 4)
    sc = _manifold(oracle, [dict(x=(0, 0, 0), he=(0.2, 0.2, 0.2), m=1.0),
                            dict(x=(0, 0.399, 0), he=(0.2, 0.2, 0.2), m=1.0, q=q)])
    assert sc.m_count[0] == 8


def test_joint_rigid_link_holds_length(oracle):
    h = _scene([dict(x=(0, 2, 0), he=(0.05, 0.05, 0.05), m=0.0, static=True),
                dict(x=(0.6, 2, 0), he=(0.1, 0.1, 0.1), m=1.0)],
               joints=[(0, 1, 0.6, 0.0)])
    sc = oracle.Scene6(h, dt=1 / 60, substeps=20, iterations=2, gravity=G, floor_y=-100.0)
    sc.step(300)
    assert abs(np.linalg.norm(sc.x[1] - sc.x[0]) - 0.6) < 0.02
    assert sc.x[1, 1] < sc.x[0, 1]


def test_stack_settles_no_penetration(oracle):
    hh = 0.2
    bodies = [dict(x=(0, hh + k * (2 * hh + 0.02), 0), he=(hh, hh, hh), m=1.0) for k in range(3)]
    sc = oracle.Scene6(_scene(bodies), dt=1 / 60, substeps=20, iterations=1, gravity=G,
                       friction=0.6)
    sc.step(350)
    ys = np.sort(sc.x[:, 1])
    assert np.allclose(np.diff(ys), 2 * hh, atol=8e-3)
    assert np.abs(sc.v).max() < 0.1


def test_pair_filters(oracle):
    """Overlapping boxes pair once (j > i); static-static and same-group>0
    pairs are dropped; the hash-grid filter (one margin) keeps a subset of the
    LBVH filter (margin on both boxes) and every truly overlapping pair."""
    bodies = [dict(x=(0, 0, 0), he=(0.2, 0.2, 0.2), m=1.0),
              dict(x=(0.35, 0, 0), he=(0.2, 0.2, 0.2), m=1.0),           # overlaps 0
              dict(x=(0.0, 0.43, 0), he=(0.2, 0.2, 0.2), m=1.0),         # within 2 margins of 0
              dict(x=(5, 0, 0), he=(0.2, 0.2, 0.2), m=0.0, static=True),
              dict(x=(5.3, 0, 0), he=(0.2, 0.2, 0.2), m=0.0, static=True),  # static-static
              dict(x=(9, 0, 0), he=(0.2, 0.2, 0.2), m=1.0, group=1),
              dict(x=(9.3, 0, 0), he=(0.2, 0.2, 0.2), m=1.0, group=1)]      # same group
    sc = oracle.Scene6(_scene(bodies), dt=1 / 60, substeps=1, iterations=1, gravity=G)
    sc.broadphase_pairs()
    lbvh = set(zip(sc.pair_a[:sc.n_pairs].tolist(), sc.pair_b[:sc.n_pairs].tolist()))
    assert lbvh == {(0, 1), (0, 2), (1, 2)}
    sc.broadphase = "hashgrid"
    sc.broadphase_pairs()
    hg = set(zip(sc.pair_a[:sc.n_pairs].tolist(), sc.pair_b[:sc.n_pairs].tolist()))
    assert hg <= lbvh and (0, 1) in hg and (0, 2) not in hg


# ---- velocity pass, caps and the gyroscopic term (kernels_6dof.py:68-102,
# 541-691): closed forms -------------------------------------------------------
def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def _velocity_floor(oracle, he, m, v, lam, particle, mu=0.0, e=0.0, q=(0, 0, 0, 1), hstep=1e-3):
    """One floor velocity pass + apply for a single body with the given
    per-corner normal multipliers; returns (v, w) after."""
    iI = (0.0, 0.0, 0.0) if particle else _box_inv_inertia(he, m)
    X, Q = _f32([[0, he[1], 0]]), _f32([q])
    V, W = _f32([v]), _f32([[0, 0, 0]])
    IM, II, HE = _f32([1.0 / m]), _f32([iI]), _f32([he])
    LAM = _f32(lam)
    DV, DW = _f32(np.zeros((1, 3))), _f32(np.zeros((1, 3)))
    DVC = np.zeros(1, np.int32)
    P = oracle._P
    lb = oracle.lib6()
    lb.orc6_velocity_floor(P(X), P(Q), P(IM), P(II), P(HE), 0.0, mu, e, P(LAM), hstep, P(V),
                           P(W), P(DV), P(DW), P(DVC), 1)
    cnt = int(DVC[0])
    lb.orc6_apply_velocity(P(IM), P(DV), P(DW), P(DVC), P(V), P(W), 1)
    return V[0].astype(float), W[0].astype(float), cnt


_BOTTOM = [c for c in range(8) if not c & 2]          # corners with sy = -1


def _bottom_lam(val):
    lam = np.zeros(8)
    lam[_BOTTOM] = val
    return lam


@pytest.mark.parametrize("e", [0.0, 0.2, 0.5, 1.0])
def test_floor_restitution_point_mass_bounces_at_e_times_impact(oracle, e):
    """A point mass (no rotational inertia) hitting the floor on four corners
    at v0 leaves it at e * v0 (kernels_6dof.py:578-579: remove the inward
    normal velocity, add e times it back outward)."""
    v, w, cnt = _velocity_floor(oracle, (0.1, 0.1, 0.1), 2.0, (0, -3.0, 0), _bottom_lam(1e-6),
                                particle=True, e=e)
    assert cnt == 4
    assert v[1] == pytest.approx(e * 3.0, rel=1e-6, abs=1e-6)
    assert np.abs(v[[0, 2]]).max() < 1e-7 and np.abs(w).max() < 1e-7


@pytest.mark.parametrize("yaw", [0.0, np.pi / 6])
def test_floor_restitution_box_generalized_mass(oracle, yaw):
    """A box with inertia: each bottom corner's impulse is (1+e) v0 divided by
    its generalized inverse mass w = 1/m + (r x n).I^-1 (r x n) (Mueller 2020
    Eq. 2-3), so the averaged update is v = -v0 + (1+e) v0 (1/m) / w; the
    torques of the four corners cancel.  A yaw about the normal leaves w
    unchanged (it maps r x n to the body frame's own corner x n), which pins
    the body/world rotation order of world_invI_mul."""
    he, m, e, v0 = (0.3, 0.1, 0.2), 2.0, 0.4, 2.5
    q = _qaa((0, 1, 0), yaw) if yaw else (0, 0, 0, 1)
    v, w, _ = _velocity_floor(oracle, he, m, (0, -v0, 0), _bottom_lam(1e-6), particle=False,
                              e=e, q=q)
    iI = np.asarray(_box_inv_inertia(he, m))
    wgen = 1 / m + he[2] ** 2 * iI[0] + he[0] ** 2 * iI[2]    # r x n = (-r_z, 0, r_x)
    want = -v0 + (1 + e) * v0 * (1 / m) / wgen
    assert v[1] == pytest.approx(want, rel=2e-6)
    assert np.abs(w).max() < 1e-5


def test_floor_separating_contact_unchanged(oracle):
    """A corner moving away from the floor gets no restitution impulse."""
    v, w, cnt = _velocity_floor(oracle, (0.1, 0.1, 0.1), 1.0, (0, 1.5, 0), _bottom_lam(1e-6),
                                particle=True, e=0.5)
    assert cnt == 0 and v[1] == pytest.approx(1.5)


@pytest.mark.parametrize("lam", [2e-7, 1e-5])
def test_floor_dynamic_friction_eq30(oracle, lam):
    """Dynamic friction (Mueller 2020 Eq. 30, kernels_6dof.py:572-576): each
    touching corner removes min(h mu |lambda / h^2|, |v_t|) of the tangential
    speed; a point mass sliding at 2 m/s keeps 2 - mu |lambda| / h, or stops
    when that exceeds its speed."""
    hstep, mu, vt = 1e-3, 0.5, 2.0
    v, w, _ = _velocity_floor(oracle, (0.1, 0.1, 0.1), 1.0, (vt, 0, 0), _bottom_lam(lam),
                              particle=True, mu=mu, hstep=hstep)
    want = vt - min(hstep * mu * abs(lam / hstep ** 2), vt)
    assert v[0] == pytest.approx(want, rel=1e-5, abs=2e-6)
    assert abs(v[1]) < 1e-7 and abs(v[2]) < 1e-7


def test_sliding_point_mass_stops_after_v0_squared_over_2_mu_g(oracle):
    """A 1 kg point-mass box sliding on the floor at v0 decelerates at mu g
    (its floor multipliers carry m g h^2 per corner) and stops after
    v0^2 / (2 mu g), the textbook Coulomb slide (15 substeps of 1/60 s)."""
    v0, mu, g = 2.0, 0.4, 9.81
    h = _scene([dict(x=(0, 0.1, 0), he=(0.1, 0.1, 0.1), m=1.0, particle=True, v=(v0, 0, 0))])
    sc = oracle.Scene6(h, dt=1 / 60, substeps=15, iterations=1, gravity=(0, -g, 0),
                       friction=mu)
    sc.step(60)
    assert abs(sc.v[0, 0]) < 1e-3
    assert sc.x[0, 0] == pytest.approx(v0 ** 2 / (2 * mu * g), rel=0.01)


def _box_pair_velocity(oracle, im_ref, v_inc, v_ref, e, mu=0.0, lam=1e-6):
    """Velocity pass of one hand-built box-box manifold: body 1 (incident)
    rests on body 0 (reference) with four points, normal +y."""
    he = (0.2, 0.2, 0.2)
    X = _f32([[0, 0, 0], [0, 0.4, 0]])
    Q = _f32([[0, 0, 0, 1], [0, 0, 0, 1]])
    V, W = _f32([v_ref, v_inc]), _f32(np.zeros((2, 3)))
    IM, II = _f32([im_ref, 1.0]), _f32(np.zeros((2, 3)))     # point masses
    MC = np.zeros(1, np.int32)
    MC[0] = 4
    MI, MR = np.ones(1, np.int32), np.zeros(1, np.int32)
    MN = _f32([[0, 1, 0]])
    corners = np.array([[sx * he[0], 0, sz * he[2]] for sx in (-1, 1) for sz in (-1, 1)])
    OI = _f32(np.zeros((8, 3)))
    OR = _f32(np.zeros((8, 3)))
    OI[:4] = corners - [0, he[1], 0]
    OR[:4] = corners + [0, he[1], 0]
    LAM = _f32(np.r_[np.full(4, lam), np.zeros(4)])
    DV, DW = _f32(np.zeros((2, 3))), _f32(np.zeros((2, 3)))
    DVC = np.zeros(2, np.int32)
    P = oracle._P
    lb = oracle.lib6()
    lb.orc6_velocity_box(P(X), P(Q), P(IM), P(II), P(MC), P(MI), P(MR), P(MN), P(OI), P(OR),
                         mu, e, P(LAM), 1e-3, P(V), P(W), P(DV), P(DW), P(DVC), 1)
    lb.orc6_apply_velocity(P(IM), P(DV), P(DW), P(DVC), P(V), P(W), 2)
    return V.astype(float)


@pytest.mark.parametrize("e", [0.0, 0.3, 1.0])
def test_box_box_restitution_static_reference(oracle, e):
    """Incident point mass landing at v0 on a static box leaves at e v0."""
    v = _box_pair_velocity(oracle, 0.0, (0, -2.0, 0), (0, 0, 0), e)
    assert v[1, 1] == pytest.approx(e * 2.0, rel=1e-6, abs=1e-6)
    assert np.abs(v[0]).max() == 0.0


@pytest.mark.parametrize("e", [0.0, 0.5])
def test_box_box_restitution_equal_masses(oracle, e):
    """Two equal point masses approaching at 2 m/s: the impulse splits by
    the generalized masses, momentum is conserved and the separation speed
    is e times the approach speed."""
    v = _box_pair_velocity(oracle, 1.0, (0, -1.5, 0), (0, 0.5, 0), e)
    assert v[0, 1] + v[1, 1] == pytest.approx(-1.0, abs=1e-6)
    assert v[1, 1] - v[0, 1] == pytest.approx(e * 2.0, abs=1e-6)


def test_box_box_dynamic_friction(oracle):
    """Tangential relative sliding between two equal point masses: each
    point removes min(h mu |lambda| / h^2, |v_t|) of the relative speed,
    split evenly, momentum conserved."""
    hstep, mu, lam = 1e-3, 0.5, 1e-6
    v = _box_pair_velocity(oracle, 1.0, (1.0, 0, 0), (-1.0, 0, 0), 0.0, mu=mu, lam=lam)
    dv = min(hstep * mu * lam / hstep ** 2, 2.0)
    assert v[1, 0] - v[0, 0] == pytest.approx(2.0 - dv, rel=1e-5)
    assert v[0, 0] + v[1, 0] == pytest.approx(0.0, abs=1e-6)


def test_cap_velocity_clamps_magnitude_keeps_direction(oracle):
    """cap_velocity_6dof: |v| > max -> v max/|v|; |w| likewise; below the
    cap and with a cap of 0 nothing changes."""
    V = _f32([[30, 40, 0], [1, 2, 2], [300, 0, 0]])
    W = _f32([[0, 0, 80], [0, 3, 4], [0, 0, 0]])
    P = oracle._P
    oracle.lib6().orc6_cap_velocity(P(V), P(W), 10.0, 20.0, 3)
    np.testing.assert_allclose(V[0], [6, 8, 0], rtol=1e-6)
    np.testing.assert_allclose(V[1], [1, 2, 2])
    np.testing.assert_allclose(W[0], [0, 0, 20], rtol=1e-6)
    np.testing.assert_allclose(W[1], [0, 3, 4])
    np.testing.assert_allclose(V[2], [10, 0, 0], rtol=1e-6)
    V2, W2 = _f32([[300, 0, 0]]), _f32([[0, 900, 0]])
    oracle.lib6().orc6_cap_velocity(P(V2), P(W2), 0.0, 0.0, 1)
    assert V2[0, 0] == 300 and W2[0, 1] == 900


def _spin(oracle, gyro, w_body, hstep, t):
    """Torque-free asymmetric box, random orientation: world angular
    momentum R I R^T w and kinetic energy, relative change after t."""
    he, m = np.array([0.3, 0.2, 0.1]), 1.0
    I = 1.0 / np.asarray(_box_inv_inertia(he, m))
    rng = np.random.default_rng(3)
    q0 = rng.normal(size=4)
    q0 /= np.linalg.norm(q0)
    X, V = _f32(np.zeros((1, 3))), _f32(np.zeros((1, 3)))
    Q = _f32([q0])
    W = _f32([_rot(q0) @ np.asarray(w_body)])
    IM, II = _f32([1.0 / m]), _f32([1.0 / I])
    XP, QP = _f32(np.zeros((1, 3))), _f32(np.zeros((1, 4)))

    def mom():
        R = _rot(Q[0].astype(float))
        Iw = R @ np.diag(I) @ R.T
        w = W[0].astype(float)
        return Iw @ w, 0.5 * w @ Iw @ w

    L0, K0 = mom()
    P = oracle._P
    lb = oracle.lib6()
    h32 = float(np.float32(hstep))
    for _ in range(int(round(t / hstep))):
        lb.orc6_integrate(P(X), P(Q), P(V), P(W), P(IM), P(II), 0.0, 0.0, 0.0, h32, gyro,
                          P(XP), P(QP), 1)
    L1, K1 = mom()
    return np.linalg.norm(L1 - L0) / np.linalg.norm(L0), abs(K1 - K0) / K0


@pytest.mark.parametrize("w_body", [(2.0, 0.5, 6.0), (6.0, 0.5, 2.0)])
def test_gyroscopic_term_conserves_world_angular_momentum(oracle, w_body):
    """With gyroscopic=True (kernels_6dof.py:91-101: w += h R I^-1 (-(w_b x
    I w_b)) in the body frame) a torque-free asymmetric box conserves its
    world angular momentum and kinetic energy up to the first-order error of
    the explicit step: the drift halves when h halves.  Without the term the
    angular velocity stays fixed in the world and the momentum is off by tens
    of percent after half a second."""
    d1, k1 = _spin(oracle, 1, w_body, 1 / 1000, 0.5)
    d2, k2 = _spin(oracle, 1, w_body, 1 / 2000, 0.5)
    d0, _ = _spin(oracle, 0, w_body, 1 / 1000, 0.5)
    assert d1 < 3e-3 and k1 < 3e-3
    assert 1.8 < d1 / d2 < 2.2 and 1.8 < k1 / k2 < 2.2
    assert d0 > 0.1
