"""Point-evaluation velocity (getVi, interaction.cpp:453-476) on the GPU
against the oracle: FP64, 1e-12 normwise."""
import numpy as np
import pytest

from conftest import normwise
from pawan_amd import capi, workloads as W

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def test_points_around_a_ring(oracle):
    s = W.ring_1k()
    rng = np.random.default_rng(6)
    pts = np.concatenate([rng.uniform(-1.5, 1.5, size=(300, 3)), s.pos[::7] + 1e-3, s.pos[:50]])
    _say(f"ring_1k: {len(pts)} points")
    ref = oracle.velocity_at(s, pts)
    with capi.Solver(s) as g:
        u = g.velocity_at(pts)
        u0, _ = g.solve("fp64")
        up = g.velocity_at(s.pos)
    assert normwise(u, ref) < 1e-12
    # one radius: at the particles the point velocity is the pair velocity
    assert normwise(up, u0) < 1e-12


def test_points_among_independent_wakes_of_many_radii(oracle):
    s = W.batched_rings(nwakes=6, nmin=64, nmax=700, seed=11)
    rng = np.random.default_rng(7)
    pts = s.pos[rng.choice(s.n, size=200, replace=False)] + rng.normal(scale=0.05, size=(200, 3))
    ref = oracle.velocity_at(s, pts)   # every particle of every wake
    with capi.Solver(s) as g:
        u = g.velocity_at(pts)
    _say(f"batched 6 wakes N={s.n}: {normwise(u, ref):.2e}")
    assert normwise(u, ref) < 1e-12


def test_points_in_the_helical_wake(oracle):
    s = W.helix_wake(1 << 20)
    rng = np.random.default_rng(8)
    pts = s.pos[rng.choice(s.n, size=16, replace=False)] + rng.normal(scale=0.01, size=(16, 3))
    ref = oracle.velocity_at(s, pts)
    with capi.Solver(s) as g:
        u = g.velocity_at(pts)
        assert g.velocity_at(np.zeros((0, 3))).shape == (0, 3)
    _say(f"helix 1M, 16 points: {normwise(u, ref):.2e}")
    assert normwise(u, ref) < 1e-12


def test_bad_arguments_are_errors():
    s = W.ring_1k()
    with capi.Solver(s) as g:
        lib = capi.load()
        u = np.zeros((2, 3))
        assert lib.pw_velocity_at(g._h, np.zeros((2, 3)), -1, u) != 0
        assert "m must be" in lib.pw_last_error().decode()
        assert lib.pw_velocity_at(g._h, np.zeros((2, 3)), 2, u) == 0
