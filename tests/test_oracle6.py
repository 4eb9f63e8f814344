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
