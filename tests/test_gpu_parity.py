"""Parity of the HIP path (through the C-ABI) with the CPU oracle.

Tolerances: FP64 paths 1e-12 normwise relative (max over particles of the
vector error / max over particles of the reference magnitude) on velocity
and vorticity rate; FP32-compensated path 1e-6 on velocity and 1e-5 on the
vorticity rate (whose magnitude is dominated by cancelling stretching
terms), both against the FP64 oracle.
"""
import numpy as np
import pytest

from conftest import normwise
from pawan_amd import capi, workloads as W

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12


def _say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("prec", ["fp64", "fp64_ref"])
def test_ring_rhs_fp64(oracle, prec):
    s = W.ring_1k()
    _say(f"ring_1k rhs {prec}")
    ref_u, ref_w = oracle.solve(s)
    with capi.Solver(s) as g:
        u, w = g.solve(prec)
    assert normwise(u, ref_u) < FP64_TOL
    assert normwise(w, ref_w) < FP64_TOL


@pytest.mark.parametrize("prec,feed", [("fp64", False), ("fp64_ref", False), ("fp64", True)])
def test_leapfrog_rhs_fp64(oracle, prec, feed):
    s = W.leapfrog_64k()
    _say(f"leapfrog_64k rhs {prec}{' scalar-fed' if feed else ''} (two coupled wakes, nu=1e-4)")
    idx = np.concatenate([np.arange(0, 256), np.arange(32768, 33024),
                          np.arange(1000, s.n, 997)])
    ref_u, ref_w = oracle.rates_targets(s, idx)
    with capi.Solver(s) as g:
        if feed:
            g.set_pair_config(prec, 8, True, False, True)
        u, w = g.solve(prec)
    assert normwise(u[idx], ref_u) < FP64_TOL
    assert normwise(w[idx], ref_w) < FP64_TOL


def test_fp32c_rhs(oracle):
    for s in (W.ring_1k(), W.leapfrog_64k()):
        _say(f"{s.name} rhs fp32c")
        idx = np.arange(0, s.n, max(1, s.n // 700))
        ref_u, ref_w = oracle.rates_targets(s, idx)
        with capi.Solver(s) as g:
            u, w = g.solve("fp32c")
        eu, ew = normwise(u[idx], ref_u), normwise(w[idx], ref_w)
        _say(f"  e_u={eu:.3e} e_w={ew:.3e}")
        assert eu < 1e-6
        assert ew < 1e-5


@pytest.mark.parametrize("scheme", ["rk4", "rk3", "euler"])
def test_ring_steps_fp64(oracle, scheme):
    s = W.ring_1k()
    _say(f"ring_1k 5 x {scheme}")
    ref = oracle.step(s, scheme, 0.005, 5)
    with capi.Solver(s) as g:
        g.step(scheme, 0.005, "fp64", 5)
        pos, alpha = g.state()
    assert normwise(pos, ref.pos) < FP64_TOL
    assert normwise(alpha, ref.alpha) < FP64_TOL


@pytest.mark.parametrize("prec,feed", [("fp64", False), ("fp64_ref", False), ("fp64", True)])
def test_independent_wakes(oracle, prec, feed):
    """Thin rings far from the origin: each wake against the oracle on that
    wake alone.  On these rings the vorticity rate is a sum of nearly
    cancelling stretching terms (a_t x a_s of almost parallel strengths), and
    the FP64 oracle itself is up to ~5e-12 (normwise) away from the
    __float128 result; the GPU's rate must be within 4x that own error of
    both the oracle and the __float128 truth (and never looser than 1e-12).
    Velocities carry no such cancellation and are held to 1e-12."""
    s = W.batched_rings(nwakes=6, nmin=64, nmax=700, seed=11)
    _say(f"batched 6 wakes N={s.n} {prec}{' scalar-fed' if feed else ''} + fp32c")
    with capi.Solver(s) as g:
        if feed:
            g.set_pair_config(prec, 2, False, False, True)
        u, w = g.solve(prec)
        u32, w32 = g.solve("fp32c")
    for k in range(s.nw):
        a, b = s.off[k], s.off[k + 1]
        ru, rw = oracle.solve(s.wake(k))
        qu, qw = oracle.solve_q128(s.wake(k))
        own = normwise(rw, qw)
        tol_w = max(FP64_TOL, 4.0 * own)
        _say(f"  wake {k} n={b - a}: e_u={normwise(u[a:b], ru):.2e} "
             f"e_w(oracle)={normwise(w[a:b], rw):.2e} e_w(q128)={normwise(w[a:b], qw):.2e} "
             f"oracle own={own:.2e}")
        assert normwise(u[a:b], ru) < FP64_TOL
        assert normwise(w[a:b], rw) < tol_w
        assert normwise(w[a:b], qw) < tol_w
        assert normwise(u32[a:b], ru) < 1e-6


def diag_scales(s, u):
    """Magnitude of the terms each diagnostic sums (a conserved quantity
    that is ~0, such as the total vorticity of a ring, is compared at the
    size of its terms, not of its rounding noise)."""
    a = np.linalg.norm(s.alpha, axis=1)
    r = np.linalg.norm(s.pos, axis=1)
    h = np.sum(a * np.linalg.norm(u, axis=1))  # helicity = -sum a.u
    return {"O": np.sum(a), "I": np.sum(r * a), "A": np.sum(r * r * a), "H": h,
            "Zc": np.max(np.abs(s.pos[:, 2]))}


def test_diagnostics_fp64(oracle):
    for s in (W.ring_1k(), W.leapfrog_64k(nst=128)):
        _say(f"{s.name} diagnostics")
        ref = oracle.diagnose(s)
        u, _ = oracle.solve(s)
        sc = diag_scales(s, u)
        with capi.Solver(s) as g:
            d = g.diagnose()[:15]
        np.testing.assert_allclose(d[0:3], ref[0:3], rtol=0, atol=1e-13 * sc["O"])
        np.testing.assert_allclose(d[3:6], ref[3:6], rtol=0, atol=1e-13 * sc["I"])
        np.testing.assert_allclose(d[6:9], ref[6:9], rtol=0, atol=1e-13 * sc["A"])
        assert abs(d[9] - ref[9]) <= 1e-12 * sc["H"]
        for m in (12, 13):  # KE, KEF (positive sums)
            assert abs(d[m] - ref[m]) <= 1e-12 * abs(ref[m]), (m, d[m], ref[m])
        assert abs(d[14] - ref[14]) <= 1e-13 * sc["Zc"], (d[14], ref[14])
        assert d[10] == 0.0 and d[11] == 0.0


def test_state_vector_layout(oracle):
    s = W.leapfrog_64k(nst=32)
    with capi.Solver() as g:
        g.set_nu(s.nu)
        for k in range(s.nw):
            a, b = s.off[k], s.off[k + 1]
            g.add_wake(s.pos[a:b], s.alpha[a:b], s.radius[a:b], s.volume[a:b], maxn=(b - a) + 17)
        st = g.get_states()
        assert len(st) == 6 * (s.n + 2 * 17)
        # wake 0 block: [pos (maxn x 3) ; alpha (maxn x 3)]
        m0 = s.off[1] + 17
        np.testing.assert_array_equal(st[:3 * s.off[1]], s.pos[:s.off[1]].ravel())
        np.testing.assert_array_equal(st[3 * m0:3 * m0 + 3 * s.off[1]], s.alpha[:s.off[1]].ravel())
        assert np.all(st[3 * s.off[1]:3 * m0] == 0)
        st2 = st.copy()
        st2[0] += 0.25
        g.set_states(st2)
        np.testing.assert_array_equal(g.get_states(), st2)
        g.solve("fp64")
        r = g.get_rates()
        u, w = oracle.solve(_shift(s, 0.25))
        np.testing.assert_allclose(r[:3 * s.off[1]].reshape(-1, 3), u[:s.off[1]], rtol=0,
                                   atol=1e-12 * np.abs(u).max())


def _shift(s, dx):
    t = s.copy()
    t.pos[0, 0] += dx
    return t
