"""Pins the CPU oracle (oracle/oracle3.c) against what the mathematics fixes.

The reference ships no golden vectors (its tests are analytic invariants,
tests/test_solver.py), so every oracle function is checked here against a
closed form, a textbook case or an independent brute force, each chosen so a
plausible slip (dropped term, wrong sign, wrong index, swapped operand) fails.
"""
import numpy as np
import pytest

G = (0.0, -9.81, 0.0)


def _one(x, m=1.0):
    x = np.asarray([x], np.float32).reshape(1, 3)
    return x, np.zeros_like(x), np.array([0.0 if m == 0 else 1.0 / m], np.float32)


def test_free_fall_discrete_closed_form(oracle):
    """integrate + finalize over k substeps: v_k = k g h and
    y_k = y0 + g h^2 k(k+1)/2 (symplectic Euler).  The FP32 finite difference
    (x - x_prev)/h carries ~ulp(x)/h of rounding per substep, a random walk of
    ~5e-3 m/s after 600 substeps, hence the tolerances."""
    x, v, w = _one((0.0, 1.0, 0.0))
    xp = np.zeros_like(x)
    h = np.float32(1.0 / 600.0)
    for _ in range(600):
        oracle.integrate(x, v, w, G, h, xp)
        oracle.finalize_velocity(x, xp, w, h, 0.0, v)
    k = 600
    hd = float(h)
    vy = k * -9.81 * hd
    y = 1.0 + -9.81 * hd * hd * k * (k + 1) / 2
    assert abs(v[0, 1] - vy) < 2e-2
    assert abs(x[0, 1] - y) < 2e-3
    assert x[0, 0] == 0.0 and x[0, 2] == 0.0


def test_static_body_and_damping(oracle):
    x, v, w = _one((1.0, 2.0, 3.0), m=0.0)
    v[:] = 5.0
    xp = np.zeros_like(x)
    oracle.integrate(x, v, w, G, 0.01, xp)
    assert (x == [[1.0, 2.0, 3.0]]).all() and (xp == x).all()
    oracle.finalize_velocity(x, xp, w, 0.01, 0.0, v)
    assert (v == 0.0).all()
    # damping multiplies the finite-difference velocity by (1 - damping)
    x, v, w = _one((0.0, 0.0, 0.0))
    xp = x.copy()
    x[0, 0] = np.float32(0.5)
    oracle.finalize_velocity(x, xp, w, 0.5, 0.25, v)
    assert v[0, 0] == np.float32(0.75)


def test_cap_velocity_direction_and_magnitude(oracle):
    v = np.array([[3.0, 4.0, 0.0], [0.3, 0.4, 0.0]], np.float32)
    oracle.cap_velocity(v, 2.5)
    assert np.allclose(v[0], [1.5, 2.0, 0.0], rtol=1e-6)
    assert (v[1] == np.float32([0.3, 0.4, 0.0])).all()


@pytest.mark.parametrize("alpha", [0.0, 1e-4])
def test_single_distance_row_closed_form(oracle, alpha):
    """One projection from lambda=0: dlam = -C/(wa+wb+alpha/h^2) and the new
    length error C' = C * alpha~/(wa+wb+alpha~); alpha=0 restores rest exactly."""
    x = np.array([[0.0, 0.0, 0.0], [1.3, 0.0, 0.0]], np.float32)
    xp = x.copy()
    w = np.array([1.0, 0.5], np.float32)
    h = 0.01
    r = oracle.rows([0], [0], [1], rest=[1.0], compliance=[alpha])
    oracle.solve_color(x, xp, w, r, [0], 0, 1, h)
    at = alpha / (h * h)
    C = 0.3
    dlam = -C / (1.5 + at)
    assert r["lam"][0] == pytest.approx(dlam, rel=1e-5)
    L = float(np.linalg.norm(x[1] - x[0]))
    assert L - 1.0 == pytest.approx(C * at / (1.5 + at), abs=2e-6)
    # the heavier body moves less: displacements are in the ratio wa : wb
    assert abs(x[0, 0]) == pytest.approx(2 * abs(x[1, 0] - 1.3), rel=1e-4)
    assert x[0, 0] > 0 and x[1, 0] < 1.3  # sign: the stretched link contracts


def test_attach_pulls_toward_anchor(oracle):
    x = np.array([[0.0, 0.0, 2.0]], np.float32)
    xp = x.copy()
    w = np.array([1.0], np.float32)
    r = oracle.rows([1], [0], [-1], rest=[0.5], anchor=[[0.0, 0.0, 0.0]])
    oracle.solve_color(x, xp, w, r, [0], 0, 1, 0.01)
    assert x[0, 2] == pytest.approx(0.5, abs=1e-6)


def test_compliance_steady_state_iteration_independent(oracle):
    """XPBD headline (reference tests/test_solver.py:50-64): the stretch of a
    compliant link holding 1 kg settles at m g alpha for any iteration count."""
    alpha = 1e-3
    for iters in (1, 4, 16):
        h = dict(x=np.array([[0, 2, 0], [0, 1, 0]], np.float32), v=np.zeros((2, 3), np.float32),
                 inv_mass=np.array([0.0, 1.0], np.float32), type=[0], a=[0], b=[1],
                 rest=[1.0], compliance=[alpha], anchor=np.zeros((1, 3)), mu=[0.0])
        sc = oracle.Scene3(h, dt=1 / 60, substeps=20, iterations=iters, gravity=G, mode="gs",
                           damping=0.05)
        sc.step(800)
        stretch = float(np.linalg.norm(sc.x[1] - sc.x[0])) - 1.0
        assert abs(stretch - 9.81 * alpha) < 0.1 * 9.81 * alpha, iters


def _floor_case(oracle, slide, mu, depth=0.01):
    """Particle at y=-depth (penetrating), displaced tangentially by `slide`
    since substep start."""
    x = np.array([[slide, -depth, 0.0]], np.float32)
    xp = np.array([[0.0, -depth, 0.0]], np.float32)
    w = np.array([1.0], np.float32)
    r = oracle.rows([2], [0], [-1], anchor=[[0.0, 0.0, 0.0]], mu=[mu])
    return x, xp, w, r


def test_floor_push_and_friction_cone(oracle):
    # inside the cone (slide < mu*d): static friction cancels the whole slide
    x, xp, w, r = _floor_case(oracle, slide=0.003, mu=0.5)
    oracle.solve_color(x, xp, w, r, [0], 0, 1, 0.01)
    assert x[0, 1] == np.float32(0.0)                 # pushed exactly to the floor
    assert abs(x[0, 0]) < 1e-8                         # slide cancelled
    assert r["lam"][0] == pytest.approx(0.01, rel=1e-6)
    # outside the cone: only mu*d of the slide is removed
    x, xp, w, r = _floor_case(oracle, slide=0.02, mu=0.5)
    oracle.solve_color(x, xp, w, r, [0], 0, 1, 0.01)
    assert x[0, 0] == pytest.approx(0.02 - 0.5 * 0.01, rel=1e-5)
    # separated rows do nothing
    x = np.array([[0.0, 0.2, 0.0]], np.float32)
    r = oracle.rows([2], [0], [-1], mu=[0.5])
    oracle.solve_color(x, x.copy(), np.array([1.0], np.float32), r, [0], 0, 1, 0.01)
    assert x[0, 1] == np.float32(0.2) and r["lam"][0] == 0.0


def test_contact_push_out_and_friction_split(oracle):
    """Overlapping spheres are pushed to rest = r_a + r_b, split by inverse
    mass; tangential friction is split by w/wsum."""
    x = np.array([[0.0, 0.0, 0.0], [0.3, 0.0, 0.0]], np.float32)
    xp = x.copy()
    w = np.array([1.0, 3.0], np.float32)
    r = oracle.rows([3], [0], [1], rest=[0.4], mu=[0.0])
    oracle.solve_color(x, xp, w, r, [0], 0, 1, 0.01)
    assert float(np.linalg.norm(x[1] - x[0])) == pytest.approx(0.4, abs=1e-6)
    assert -x[0, 0] == pytest.approx((x[1, 0] - 0.3) / 3.0, rel=1e-5)
    # friction inside the cone cancels the relative tangential slide
    x = np.array([[0.0, 0.0, 0.0], [0.3, 0.002, 0.0]], np.float32)
    xp = np.array([[0.0, 0.0, 0.0], [0.3, 0.0, 0.0]], np.float32)
    n = (x[0] - x[1]) / np.linalg.norm(x[0] - x[1])   # normal at projection time
    r = oracle.rows([3], [0], [1], rest=[0.4], mu=[1.0])
    oracle.solve_color(x, xp, w, r, [0], 0, 1, 0.01)
    dp = (x[0] - xp[0]) - (x[1] - xp[1])
    assert np.linalg.norm(dp - n * np.dot(dp, n)) < 5e-7


@pytest.mark.parametrize("friction", [False, True])
def test_jacobi_equals_gs_on_one_color_class(oracle, friction):
    """Rows that share no particle form one color; with relax=1, Jacobi's
    averaged apply (count 1 per particle) equals the Gauss-Seidel sweep.
    Without friction the two are the same float operations (bit-exact); with
    friction GS adds the normal and friction corrections to x one after the
    other while Jacobi sums them first, so they agree to rounding."""
    rng = np.random.default_rng(3)
    n = 64
    x = rng.normal(size=(n, 3)).astype(np.float32)
    xp = (x + rng.normal(scale=1e-3, size=x.shape)).astype(np.float32)
    w = rng.uniform(0.2, 2.0, n).astype(np.float32)
    w[5] = 0.0
    perm = rng.permutation(n)
    a, b = perm[0::2][:24], perm[1::2][:24]
    types = np.array([0] * 8 + [3] * 8 + [0] * 8, np.int32)
    rest = rng.uniform(0.5, 3.0, 24).astype(np.float32)
    comp = np.where(types == 0, 1e-5, 0.0).astype(np.float32)
    mu = np.where(types == 3, 0.4 if friction else 0.0, 0.0).astype(np.float32)
    # one-body rows (FLOOR, ATTACH) on particles not used above
    fl = perm[48:56]
    types = np.concatenate([types, np.array([2] * 4 + [1] * 4, np.int32)])
    a = np.concatenate([a, fl])
    b = np.concatenate([b, np.full(8, -1)])
    rest = np.concatenate([rest, np.zeros(4, np.float32), np.full(4, 0.3, np.float32)])
    comp = np.concatenate([comp, np.zeros(4, np.float32), np.full(4, 1e-6, np.float32)])
    anc = np.zeros((32, 3), np.float32)
    anc[24:28, 1] = 0.5
    anc[28:] = rng.normal(size=(4, 3))
    mu = np.concatenate([mu, np.full(4, 0.3 if friction else 0.0, np.float32),
                         np.zeros(4, np.float32)])
    r1 = oracle.rows(types, a, b, rest, comp, anc, mu)
    r2 = oracle.rows(types, a, b, rest, comp, anc, mu)
    x1, x2 = x.copy(), x.copy()
    oracle.solve_color(x1, xp, w, r1, np.arange(32), 0, 32, 1 / 600)
    dx = np.zeros_like(x)
    dn = np.zeros(n, np.int32)
    oracle.solve_jacobi(x2, xp, w, r2, 1 / 600, dx, dn)
    oracle.apply_jacobi(x2, dx, dn, w, 1.0)
    assert (r1["lam"] == r2["lam"]).all()
    assert (r1["lam"] != 0).sum() >= 16
    if friction:
        np.testing.assert_allclose(x1, x2, rtol=1e-6, atol=1e-7)
        assert not (x1 == x).all()
    else:
        assert (x1 == x2).all()


def _greedy_certificate(n_bodies, a, b):
    """Independent brute force of the reference greedy: visit rows by a
    stable sort on -(deg a + deg b), give each the mex of the colors of
    already-colored rows sharing a body."""
    m = len(a)
    deg = [0] * n_bodies
    for j in range(m):
        deg[a[j]] += 1
        if b[j] >= 0:
            deg[b[j]] += 1
    keyed = sorted(range(m), key=lambda j: -(deg[a[j]] + (deg[b[j]] if b[j] >= 0 else 0)))
    color = [-1] * m
    for j in keyed:
        ends = {a[j]} | ({b[j]} if b[j] >= 0 else set())
        taken = {color[k] for k in range(m)
                 if color[k] >= 0 and ({a[k]} | ({b[k]} if b[k] >= 0 else set())) & ends}
        c = 0
        while c in taken:
            c += 1
        color[j] = c
    return np.array(color, np.int32)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_coloring_matches_greedy_certificate(oracle, seed):
    rng = np.random.default_rng(seed)
    n_bodies = 40
    a = rng.integers(0, n_bodies, 200)
    b = rng.integers(0, n_bodies, 200)
    b = np.where(b == a, (b + 1) % n_bodies, b)      # reference test_solver.py:122-138
    b[::7] = -1                                       # one-body rows
    got = oracle.color_constraints(n_bodies, a, b)
    assert (got == _greedy_certificate(n_bodies, a.tolist(), b.tolist())).all()
    # validity: no two same-color rows share a body
    for body in range(n_bodies):
        rows_at = [j for j in range(200) if a[j] == body or b[j] == body]
        cols = [int(got[j]) for j in rows_at]
        assert len(cols) == len(set(cols))


def test_color_ranges_is_stable_csr(oracle):
    color = np.array([2, 0, 1, 0, 2, 1, 0], np.int32)
    order, starts, k = oracle.color_ranges(color)
    assert k == 3
    assert order.tolist() == [1, 3, 6, 2, 5, 0, 4]
    assert starts.tolist() == [0, 3, 5, 7]
