"""The committed fixtures of tests/golden/ (written by make_golden.py) are
reproduced by the generators and the oracle on this machine, and by the HIP
path on the GPU."""
import os

import numpy as np
import pytest

from conftest import normwise
from pawan_amd import workloads as W

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return dict(np.load(os.path.join(HERE, f"{name}.npz"), allow_pickle=False))


SYSTEMS = {
    "ring_1k": W.ring_1k,
    "leapfrog_64k": W.leapfrog_64k,
    "leapfrog_256_rk3": lambda: W.leapfrog_64k(nst=16),
    "batched_6": lambda: W.batched_rings(nwakes=6, nmin=64, nmax=700, seed=11),
    "helix_1m": lambda: W.helix_wake(1 << 20),
}


@pytest.mark.parametrize("name", sorted(SYSTEMS))
def test_generator_reproduces_fixture_inputs(name):
    g = load(name)
    s = SYSTEMS[name]()
    rows = g["in_rows"]
    np.testing.assert_array_equal(s.off, g["in_off"])
    assert s.nu == g["in_nu"]
    np.testing.assert_allclose(s.pos[rows], g["in_pos"], rtol=0, atol=4e-16 * np.abs(g["in_pos"]).max())
    np.testing.assert_allclose(s.alpha[rows], g["in_alpha"], rtol=4e-15, atol=4e-16 * np.abs(g["in_alpha"]).max())
    np.testing.assert_allclose(s.radius[rows], g["in_radius"], rtol=1e-15)
    np.testing.assert_allclose(s.volume[rows], g["in_volume"], rtol=4e-15)


def test_oracle_reproduces_ring(oracle):
    g = load("ring_1k")
    s = W.ring_1k()
    u, w = oracle.solve(s)
    assert normwise(u, g["vel"]) < 1e-14 and normwise(w, g["retvor"]) < 1e-14
    st = oracle.step(s, "rk4", 0.005, 2)
    assert normwise(st.pos, g["rk4x2_pos"]) < 1e-15
    d = oracle.diagnose(s)
    np.testing.assert_allclose(d, g["diag"], rtol=1e-13, atol=1e-16)


def test_oracle_reproduces_sampled_targets(oracle):
    for name in ("leapfrog_64k",):
        g = load(name)
        s = SYSTEMS[name]()
        u, w = oracle.rates_targets(s, g["targets"])
        assert normwise(u, g["vel"]) < 1e-14 and normwise(w, g["retvor"]) < 1e-14


def test_oracle_reproduces_rk3_steps(oracle):
    g = load("leapfrog_256_rk3")
    st = oracle.step(SYSTEMS["leapfrog_256_rk3"](), "rk3", 0.01, 3)
    assert normwise(st.pos, g["pos"]) < 1e-14 and normwise(st.alpha, g["alpha"]) < 1e-13


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["ring_1k", "leapfrog_64k", "helix_1m"])
@pytest.mark.parametrize("prec,tol_u,feed", [("fp64", 1e-12, False), ("fp64", 1e-12, True), ("fp32c", 1e-6, False),
                                             ("fp32c", 1e-6, True)])
def test_gpu_matches_fixture(name, prec, tol_u, feed):
    from pawan_amd import capi

    g = load(name)
    s = SYSTEMS[name]()
    print(f"{name} N={s.n} {prec}{' scalar-fed' if feed else ''} vs fixture", flush=True)
    with capi.Solver(s) as solver:
        if feed:
            solver.set_pair_config(prec, 4, True, prec == "fp32c", True)
        u, w = solver.solve(prec)
    idx = g["targets"] if "targets" in g else np.arange(s.n)
    eu, ew = normwise(u[idx], g["vel"]), normwise(w[idx], g["retvor"])
    print(f"  e_u={eu:.3e} e_w={ew:.3e}", flush=True)
    assert eu < tol_u
    assert ew < (1e-12 if prec == "fp64" else 1e-5)


@pytest.mark.gpu
def test_gpu_rk_steps_match_fixtures():
    from pawan_amd import capi

    g = load("ring_1k")
    with capi.Solver(W.ring_1k()) as solver:
        solver.step("rk4", 0.005, "fp64", 2)
        pos, alpha = solver.state()
        d = solver.diagnose()[:15]
    assert normwise(pos, g["rk4x2_pos"]) < 1e-12 and normwise(alpha, g["rk4x2_alpha"]) < 1e-12
    np.testing.assert_allclose(d[12:14], g["diag_rk4x2"][12:14], rtol=1e-12)
    g = load("leapfrog_256_rk3")
    with capi.Solver(SYSTEMS["leapfrog_256_rk3"]()) as solver:
        solver.step("rk3", 0.01, "fp64", 3)
        pos, alpha = solver.state()
    assert normwise(pos, g["pos"]) < 1e-12 and normwise(alpha, g["alpha"]) < 1e-12
