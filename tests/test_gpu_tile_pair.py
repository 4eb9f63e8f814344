"""The two-iteration tile launch (xpbd_jacobi_tile_pair) against the plan
of two launches per substep (xpbd_jacobi_tile + xpbd_jacobi_tile_next): the
launch re-runs the first iteration of its ring particles with the same
exact fixed-point sums, so every frame must be bit for bit the same."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _run(build, pair, frames):
    s = build()
    s.tile_pair = pair
    s.block_frames = False
    out = []
    for _ in range(frames):
        s.step()
        out.append((s.positions(), s.velocities(), s.lambdas()))
    used = s._pair_dev is not None
    return out, used


def _same(build, frames):
    ref, used_ref = _run(build, False, frames)
    got, used = _run(build, True, frames)
    assert used and not used_ref
    for f, (g, e) in enumerate(zip(got, ref)):
        for name, a, b in zip(("x", "v", "lambda"), g, e):
            assert np.array_equal(a, b), (f, name, float(np.abs(a - b).max()))
    return got


@pytest.mark.parametrize("tile", [64, 32, 16])
def test_pair_launch_equals_two_launches_with_attach_and_gravity_scale(tile):
    """36^2 cloth with per-particle gravity scales, an ATTACH row and
    FLOOR rows (8 frames); tile 32 and 16 leave rings that reach across
    several tiles."""
    from xpbd_amd import Solver
    from xpbd_amd.scenes import cloth_sheet
    from xpbd_amd.solver import Body

    def build():
        s, _ = cloth_sheet(Solver, res=36, solve_mode="jacobi", device=DEV, iterations=2)
        s.tile_size = tile
        s.set_gravity_scale(np.linspace(0.5, 1.5, s.num_bodies))
        s.add_attach(Body(5), (0.1, 0.6, 0.0), compliance=1e-6)
        return s

    _same(build, 8)


def test_pair_launch_equals_two_launches_on_the_floor_with_friction():
    """A 24^2 cloth dropped from 5 cm: after the first frames its FLOOR rows
    are active with friction (they read x_prev of ring particles), damping
    and the speed cap on."""
    from xpbd_amd import Solver
    from xpbd_amd.scenes import cloth_sheet

    def build():
        s, _ = cloth_sheet(Solver, res=24, height=0.05, solve_mode="jacobi", device=DEV,
                           iterations=2, damping=0.01, max_speed=3.0)
        s.tile_size = 32
        return s

    got = _same(build, 24)
    assert got[-1][0][:, 1].min() < 0.02  # resting on the floor


@pytest.mark.parametrize("substeps", [1, 2, 3])
def test_pair_launch_edge_shapes(substeps):
    """49 particles (one partial tile at 64, two at 32): one substep never
    reaches the pair launch, two use it once (as the last substep), three
    once in the middle."""
    from xpbd_amd import Solver
    from xpbd_amd.scenes import cloth_sheet

    for tile in (64, 32):
        def build():
            s, _ = cloth_sheet(Solver, res=7, solve_mode="jacobi", device=DEV,
                               substeps=substeps, iterations=2)
            s.tile_size = tile
            return s

        if substeps == 1:
            ref, _ = _run(build, False, 6)
            got, _ = _run(build, True, 6)
            for g, e in zip(got, ref):
                for a, b in zip(g, e):
                    assert np.array_equal(a, b)
        else:
            _same(build, 6)


def test_pair_launch_on_the_headline_cloth():
    """The 128^2 cloth, 15 substeps x 2 iterations, 6 frames, with the
    solver's default knobs (graph replay of the frame plan)."""
    from xpbd_amd import Solver
    from xpbd_amd.scenes import cloth_sheet

    out = {}
    for pair in (False, True):
        s, _ = cloth_sheet(Solver, res=128, solve_mode="jacobi", device=DEV)
        s.tile_pair = pair
        s.step(6)
        out[pair] = (s.positions(), s.velocities(), s.lambdas(), s._pair_dev is not None)
    assert out[True][3] and not out[False][3]
    for a, b in zip(out[True][:3], out[False][:3]):
        assert np.array_equal(a, b)
