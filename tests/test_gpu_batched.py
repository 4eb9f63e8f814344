"""The batched workload at full size on the GPU: 1024 uneven scenes (chains,
cloths, sphere piles with device CONTACT rows, box piles; per-scene gravity
scale, compliance and friction) in one particle system and one rigid
system, and the learned warm start on it.

* The world steps finite, piles stay on their floors, scenes stay apart.
* Sampled scenes (a chain, a cloth, a sphere pile with the device's contact
  rows, a box pile) of a one-substep copy of the world, loaded with the full
  world's state, match the oracle stepped on that scene alone from the same
  state: positions, velocities and multipliers within 1e-4 relative to the
  scene's own extent; the box pile's pair list and per-pair contact counts
  bit-exact.
* The learned warm start fitted on the world: one iteration with the network
  once per frame reaches the residual of two plain iterations.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
G = -9.81


@pytest.fixture(scope="module")
def worlds():
    from xpbd_amd.batch import BatchedWorld, draw

    spec = draw(1024, 0)
    full = BatchedWorld(spec, device=DEV)
    h = float(np.float32(full.particles.dt / full.particles.substeps))
    one = BatchedWorld(draw(1024, 0), device=DEV, substeps=1, dt=h)
    return spec, full, one


def _rel(g, o):
    g = np.asarray(g, np.float64)
    o = np.asarray(o, np.float64)
    return float(np.abs(g - o).max()) / max(float(np.abs(o).max()), 1e-30)


def test_config4_world_full_size_steps_finite(worlds):
    spec, w, _ = worlds
    tot = spec.totals()
    assert tot["scenes"] == 1024 and tot["total"] > 1_600_000
    assert {sc.kind for sc in spec.scenes} == {"chain", "cloth", "pile", "boxes"}
    for _ in range(20):
        w.step()
    x = w.particles.positions()
    assert np.all(np.isfinite(x))
    xr = w.rigid.positions()
    assert np.all(np.isfinite(xr))
    assert w.particles.n_contacts > 1000
    for sc in spec.scenes:
        if sc.kind == "pile":
            y = x[sc.first:sc.first + sc.count, 1]
            assert y.min() > 0.05 - 0.01, sc
        if sc.kind == "boxes":
            y = xr[sc.first:sc.first + sc.count, 1]
            assert y.min() > 0.16 - 0.03, sc
        if sc.kind in ("pile", "boxes"):
            # colliding scenes stay inside their own 3 m cell
            pts = (x if sc.kind == "pile" else xr)[sc.first:sc.first + sc.count]
            off = pts[:, [0, 2]] - np.asarray(sc.origin)[[0, 2]]
            assert off.min() > -0.5 and off.max() < 2.5, sc
    print(f"config4: {tot}, {w.particles.n_contacts} device contact rows, "
          f"{w.rigid.n_pairs} rigid pairs")


def _pick(spec, kind, lo, hi):
    for sc in spec.scenes:
        if sc.kind == kind and lo <= sc.size <= hi:
            return sc
    raise AssertionError(f"no {kind} scene of size {lo}-{hi}")


def _scene_rows(h, first, count, extra=None):
    """Rows of the scene's particles (both endpoints in range), re-indexed."""
    a, b = h["a"].astype(np.int64), h["b"].astype(np.int64)
    sel = (a >= first) & (a < first + count)
    out = {k: h[k][sel] for k in ("type", "rest", "compliance", "mu")}
    out["anchor"] = h["anchor"].reshape(-1, 3)[sel]
    out["a"] = (a[sel] - first).astype(np.int32)
    out["b"] = np.where(b[sel] >= 0, b[sel] - first, -1).astype(np.int32)
    assert np.all(out["b"] < count)
    return out, sel


def test_config4_sampled_scenes_match_oracle_per_substep(worlds, oracle):
    spec, w, one = worlds
    for _ in range(25):
        w.step()
    P, P1 = w.particles, one.particles
    x, v = P.positions(), P.velocities()
    P1.set_state(x, v)
    one.rigid._flush()
    for k in ("X", "Q", "V", "W"):
        one.rigid._buf[k].copy_(w.rigid._buf[k])
    one.step()
    x1, v1, lam1 = P1.positions(), P1.velocities(), P1.lambdas()
    h_all = P1.host_arrays()
    # the device's contact rows this frame, in API indices
    cnt = int(P1.n_contacts)
    crow = P1.CROWS[:cnt].cpu().numpy()
    ca = P1._pperm[(crow[:, 0] & 0x0FFFFFFF).astype(np.int64)]
    cb = P1._pperm[crow[:, 1].astype(np.int64)]
    clam = P1.CLAM[:cnt].cpu().numpy()
    hstep = P1.dt
    for kind, lo, hi in (("chain", 20, 120), ("cloth", 20, 40), ("pile", 150, 700)):
        sc = _pick(spec, kind, lo, hi)
        f, n = sc.first, sc.count
        rows, sel = _scene_rows(h_all, f, n)
        m_static = len(rows["type"])
        csel = (ca >= f) & (ca < f + n)
        if kind == "pile":
            assert csel.sum() > 50
            assert np.all((cb[csel] >= f) & (cb[csel] < f + n))
            nc = int(csel.sum())
            rows["type"] = np.r_[rows["type"], np.full(nc, 3, np.int32)]
            rows["a"] = np.r_[rows["a"], (ca[csel] - f).astype(np.int32)]
            rows["b"] = np.r_[rows["b"], (cb[csel] - f).astype(np.int32)]
            rows["rest"] = np.r_[rows["rest"], crow[csel, 2].copy().view(np.float32)]
            rows["mu"] = np.r_[rows["mu"], crow[csel, 3].copy().view(np.float32)]
            rows["compliance"] = np.r_[rows["compliance"], np.zeros(nc, np.float32)]
            rows["anchor"] = np.r_[rows["anchor"], np.zeros((nc, 3), np.float32)]
        else:
            assert csel.sum() == 0
        scene = dict(rows)
        scene["x"], scene["v"] = x[f:f + n], v[f:f + n]
        scene["inv_mass"] = h_all["inv_mass"][f:f + n]
        gy = float(np.float32(G) * np.float32(sc.gscale))
        orc = oracle.Scene3(scene, dt=hstep, substeps=1, iterations=P1.iterations,
                            gravity=(0.0, gy, 0.0), mode="jacobi", damping=P1.damping,
                            relax=P1.jacobi_relax, max_speed=P1.max_speed)
        orc.step(1)
        origin = np.asarray(sc.origin, np.float32)
        lam_dev = lam1[sel]
        if kind == "pile":
            lam_dev = np.r_[lam_dev, clam[csel]]
        errs = (_rel(x1[f:f + n] - origin, orc.x - origin), _rel(v1[f:f + n], orc.v),
                _rel(lam_dev, orc.lam))
        print(f"config4 {kind} {sc.size} ({n} particles, {len(orc.lam)} rows, "
              f"{m_static} static): x/v/lam rel err {errs}")
        assert max(errs) <= 1e-4, (kind, errs)
    # a box pile against the 6-DOF oracle on that pile alone
    sc = _pick(spec, "boxes", 20, 80)
    f, n = sc.first, sc.count
    R1 = one.rigid
    hr = R1.host_arrays()
    sub = {k: hr[k][f:f + n] for k in ("q", "inv_mass", "inv_I", "he", "group")}
    for k, t in (("x", "X"), ("v", "V"), ("w", "W")):
        sub[k] = w.rigid._buf[t][f:f + n, :3].cpu().numpy()
    sub["q"] = w.rigid._buf["Q"][f:f + n].cpu().numpy()
    gy = float(np.float32(G) * np.float32(sc.gscale))
    orc6 = oracle.Scene6(sub, dt=R1.dt, substeps=1, iterations=R1.iterations,
                         gravity=(0.0, gy, 0.0), floor_y=R1.floor_y, friction=R1.friction,
                         restitution=R1.restitution, relax=R1.jacobi_relax,
                         max_speed=R1.max_speed, max_omega=R1.max_omega)
    orc6.step(1)
    pr = R1.pairs()
    mine = (pr[:, 0] >= f) & (pr[:, 0] < f + n)
    want = np.stack([orc6.pair_a[:orc6.n_pairs], orc6.pair_b[:orc6.n_pairs]], 1) + f
    assert np.array_equal(pr[mine], want)
    assert np.array_equal(R1.contact_counts()[mine], orc6.m_count[:orc6.n_pairs])
    origin = np.asarray(sc.origin, np.float32)
    errs = (_rel(R1.positions()[f:f + n] - origin, orc6.x - origin),
            _rel(R1.orientations()[f:f + n], orc6.q), _rel(R1.velocities()[f:f + n], orc6.v))
    print(f"config4 boxes {n}: {int(mine.sum())} pairs, "
          f"{int(orc6.m_count[:orc6.n_pairs].sum())} points, x/q/v rel err {errs}")
    assert max(errs) <= 1e-4, errs


def test_config5_batched_warm_start_halves_the_iterations():
    """Config 5 on the batched world: the network fitted on the world for one
    pass per frame (two rounds on the GPU); from the same state each start
    on its own trajectory over 60 substeps; the learned start with one
    iteration reaches at most plain-2's RMS residual over the static rows,
    with two at most plain-4's."""
    from xpbd_amd import warmstart as ws
    from xpbd_amd.batch import BatchedWorld, draw

    w = BatchedWorld(draw(1024, 0), device=DEV)
    s = w.particles
    s.block_frames = False
    s.step(2)
    x0, v0 = s.positions(), s.velocities()

    def restore():
        if s._ws_theta is not None:
            s.set_warm_start(None)
        s.iterations = 2
        s.set_state(x0, v0)
        return s

    theta = ws.fit(restore, substeps=60, settle_frames=1, steps=800, net_every=15,
                   pair_with_net=True, every=5)
    restore()
    s.step(4)
    res, fb = ws.Evaluator(s, theta).run(substeps=60, plain_iters=(1, 2, 4),
                                         learned_iters=(1, 2), net_every=15)
    print("batched rms residual (network once per frame):", res)
    assert res["learned_1"] <= res["plain_2"]
    assert res["learned_2"] <= res["plain_4"]
    assert res["learned_1"] < res["plain_1"]
