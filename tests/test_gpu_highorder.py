"""The high-order algebraic smoothing kernel (highorder.h:20-106) on the GPU
against the oracle under the same kernel: right-hand sides (FP64 and
FP32C), steps, diagnostics and point velocities, and the kernel switch
itself."""
import numpy as np
import pytest

from conftest import normwise
from pawan_amd import capi, workloads as W

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12


def _say(msg):
    print(msg, flush=True)


@pytest.fixture
def ho(oracle):
    with oracle.smoothing(oracle.HIGHORDER):
        yield oracle


def _solver(s):
    g = capi.Solver(s)
    g.set_kernel(capi.PW_KERNEL_HIGHORDER)
    return g


@pytest.mark.parametrize("prec", ["fp64", "fp64_ref"])
def test_ring_rhs(ho, prec):
    s = W.ring_1k()
    _say(f"ring_1k high-order rhs {prec}")
    ref_u, ref_w = ho.solve(s)
    with _solver(s) as g:
        assert g.kernel == capi.PW_KERNEL_HIGHORDER
        u, w = g.solve(prec)
    _say(f"  e_u={normwise(u, ref_u):.2e} e_w={normwise(w, ref_w):.2e}")
    assert normwise(u, ref_u) < FP64_TOL
    assert normwise(w, ref_w) < FP64_TOL


def test_leapfrog_rhs(ho):
    s = W.leapfrog_64k()
    _say("leapfrog_64k high-order rhs (two coupled wakes, nu=1e-4)")
    idx = np.concatenate([np.arange(0, 256), np.arange(32768, 33024), np.arange(1000, s.n, 997)])
    ref_u, ref_w = ho.rates_targets(s, idx)
    with _solver(s) as g:
        g.set_pair_config("fp64", parts=4, persistent=True)
        u, w = g.solve("fp64")
    _say(f"  e_u={normwise(u[idx], ref_u):.2e} e_w={normwise(w[idx], ref_w):.2e}")
    assert normwise(u[idx], ref_u) < FP64_TOL
    assert normwise(w[idx], ref_w) < FP64_TOL


def test_independent_wakes(ho):
    """Thin rings of many radii: each wake against the oracle on that wake
    alone.  The vorticity rate there is a sum of nearly cancelling
    stretching terms; the Gaussian oracle itself sits up to ~5e-12 from a
    __float128 evaluation on these rings, so the rate is held to 2e-11."""
    s = W.batched_rings(nwakes=6, nmin=64, nmax=700, seed=11)
    _say(f"batched 6 wakes N={s.n} high-order")
    with _solver(s) as g:
        u, w = g.solve("fp64")
    for k in range(s.nw):
        a, b = s.off[k], s.off[k + 1]
        ru, rw = ho.solve(s.wake(k))
        _say(f"  wake {k} n={b - a}: e_u={normwise(u[a:b], ru):.2e} e_w={normwise(w[a:b], rw):.2e}")
        assert normwise(u[a:b], ru) < FP64_TOL
        assert normwise(w[a:b], rw) < 2e-11


def test_steps_and_kernel_switch(oracle):
    """Two Gaussian RK4 steps, then two high-order ones on the same context
    (graph-captured steps are rebuilt when the kernel changes)."""
    s = W.ring_1k()
    _say("ring_1k 2 x rk4 Gaussian + 2 x rk4 high-order, graphs on")
    mid = oracle.step(s, "rk4", 0.005, 2)
    with oracle.smoothing(oracle.HIGHORDER):
        ref = oracle.step(mid, "rk4", 0.005, 2)
    with capi.Solver(s) as g:
        g.set_graph(True)
        g.step("rk4", 0.005, "fp64", 2)
        pos, alpha = g.state()
        assert normwise(pos, mid.pos) < FP64_TOL and normwise(alpha, mid.alpha) < FP64_TOL
        g.set_kernel(capi.PW_KERNEL_HIGHORDER)
        g.step("rk4", 0.005, "fp64", 2)
        pos, alpha = g.state()
    _say(f"  e_pos={normwise(pos, ref.pos):.2e} e_alpha={normwise(alpha, ref.alpha):.2e}")
    assert normwise(pos, ref.pos) < FP64_TOL
    assert normwise(alpha, ref.alpha) < FP64_TOL


def test_diagnostics(ho):
    for s in (W.ring_1k(), W.leapfrog_64k(nst=128)):
        _say(f"{s.name} high-order diagnostics")
        ref = ho.diagnose(s)
        u, _ = ho.solve(s)
        h_scale = np.sum(np.linalg.norm(s.alpha, axis=1) * np.linalg.norm(u, axis=1))
        with _solver(s) as g:
            d = g.diagnose()[:15]
        _say(f"  H {d[9]:.6e}/{ref[9]:.6e} E {d[10]:.6e}/{ref[10]:.6e} EF {d[11]:.6e}/{ref[11]:.6e}")
        assert abs(d[9] - ref[9]) <= 1e-12 * h_scale
        for m in (10, 11, 12, 13):  # enstrophies and kinetic energies
            assert abs(d[m] - ref[m]) <= 1e-12 * abs(ref[m]), (m, d[m], ref[m])
        assert ref[10] != 0.0 and ref[11] != 0.0


def test_point_velocity(ho):
    s = W.batched_rings(nwakes=6, nmin=64, nmax=700, seed=11)
    rng = np.random.default_rng(9)
    pts = s.pos[rng.choice(s.n, size=200, replace=False)] + rng.normal(scale=0.05, size=(200, 3))
    ref = ho.velocity_at(s, pts)
    with _solver(s) as g:
        u = g.velocity_at(pts)
    _say(f"batched high-order point velocity: {normwise(u, ref):.2e}")
    assert normwise(u, ref) < FP64_TOL


@pytest.mark.parametrize("feed,t2", [(False, False), (True, False), (True, True)])
def test_fp32c_rhs(ho, feed, t2):
    """The compensated FP32 form of the kernel (LDS-tiled, scalar-fed, and
    scalar-fed with two targets per lane), at
    the Gaussian FP32C bars: velocity 1e-6 and vorticity rate 1e-5 normwise
    against the FP64 oracle."""
    for s in (W.ring_1k(), W.leapfrog_64k()):
        idx = np.arange(0, s.n, max(1, s.n // 700))
        ref_u, ref_w = ho.rates_targets(s, idx)
        with _solver(s) as g:
            g.set_pair_config("fp32c", 1, False, False, feed, t2)
            u, w = g.solve("fp32c")
            u64, w64 = g.solve("fp64")
        eu, ew = normwise(u[idx], ref_u), normwise(w[idx], ref_w)
        _say(f"{s.name} high-order fp32c: e_u={eu:.3e} e_w={ew:.3e} "
             f"(vs own fp64: {normwise(u, u64):.2e}, {normwise(w, w64):.2e})")
        assert eu < 1e-6
        assert ew < 1e-5
        assert normwise(u, u64) < 1e-6 and normwise(w, w64) < 1e-5


@pytest.mark.parametrize("feed", [False, True])
def test_fp32c_separated_tiles_on_the_helix(ho, feed):
    """On the 1M helix most tile pairs take the separated loop (FP32
    positions relative to the targets' or the source chunk's centre, the
    kernel evaluated exactly): sampled targets against the oracle."""
    s = W.helix_wake(1 << 20)
    idx = np.random.default_rng(3).choice(s.n, size=48, replace=False)
    ref_u, ref_w = ho.rates_targets(s, idx)
    with _solver(s) as g:
        g.set_pair_config("fp32c", 4, True, True, feed)
        u, w = g.solve("fp32c")
    eu, ew = normwise(u[idx], ref_u), normwise(w[idx], ref_w)
    _say(f"helix 1M high-order fp32c, 48 targets: e_u={eu:.3e} e_w={ew:.3e}")
    assert eu < 1e-6
    assert ew < 1e-5


def test_autotune_under_the_high_order_kernel(ho, tmp_path, monkeypatch):
    """The autotuner times the high-order kernels and keys its result by
    them (family 4 = FP64, 5 = FP32C in the tune file, apart from the
    Gaussian's 1 and 2); the tuned launches keep the oracle's bars (for the
    FP32C vorticity rate of these thin rings, the conditioning-scaled bar
    of the batched Gaussian test: 8 x 2^29 times the wake's FP64 error,
    1e-5 at least)."""
    path = tmp_path / "tune.txt"
    monkeypatch.setenv("PAWAN_TUNE_CACHE", str(path))
    s = W.batched_rings(nwakes=5, nmin=300, nmax=2500, seed=23)
    ref = [ho.solve(s.wake(k)) for k in range(s.nw)]
    ew64 = np.zeros(s.nw)
    with _solver(s) as g:
        for prec, fam in (("fp64", 4), ("fp32c", 5)):
            ms = g.autotune(prec)
            u, w = g.solve(prec)
            _say(f"high-order {prec} tuned: {g.pair_config(prec)} {ms:.3f} ms")
            for k in range(s.nw):
                a, b = s.off[k], s.off[k + 1]
                eu, ew = normwise(u[a:b], ref[k][0]), normwise(w[a:b], ref[k][1])
                _say(f"  wake {k} n={b - a}: e_u={eu:.2e} e_w={ew:.2e}")
                if prec == "fp64":
                    assert eu < FP64_TOL and ew < 2e-11
                    ew64[k] = ew
                else:
                    assert eu < 1e-6 and ew < max(1e-5, 8.0 * 2.0 ** 29 * ew64[k])
            lines = [l.split() for l in path.read_text().splitlines()]
            assert any(int(l[0]) == fam and int(l[1]) == s.n for l in lines), lines
        g.set_kernel(capi.PW_KERNEL_GAUSSIAN)
        g.autotune("fp32c")  # same shape, other kernel: a search of its own
    lines = [l.split() for l in path.read_text().splitlines()]
    assert sorted(int(l[0]) for l in lines if int(l[1]) == s.n) == [2, 4, 5]


def test_bad_kernel_is_an_error():
    s = W.ring_1k()
    with _solver(s) as g:
        with pytest.raises(capi.PawanError) as e:
            g.set_kernel(7)
        assert e.value.code == capi.PW_ERR_ARG
        g.set_kernel(capi.PW_KERNEL_GAUSSIAN)
        u, w = g.solve("fp32c")
    with capi.Solver(s) as g2:
        u2, w2 = g2.solve("fp32c")
    np.testing.assert_array_equal(u, u2)
    np.testing.assert_array_equal(w, w2)
