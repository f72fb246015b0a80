"""Pins of the oracle's diagnostics (interaction.cpp:64-194,
interaction_utils.h:107-331) by closed forms and identities."""
import math

import numpy as np
import pytest

from pawan_amd import workloads as W

O_, I_, A_, H, E, EF, KE, KEF, ZC = slice(0, 3), slice(3, 6), slice(6, 9), 9, 10, 11, 12, 13, 14


def test_thin_ring_impulses(oracle):
    """Thin ring of circulation Gamma, radius R (ring.cpp:22-39): total
    vorticity 0, linear impulse (0, 0, pi R^2 Gamma) exactly for the discrete
    sum, angular impulse 0 about its centre and (0, -pi Gamma R^2 c, 0) when
    moved by c along x."""
    R, G, n = 1.3, 0.7, 64
    s = W.thin_ring(R, G, 0.1, n)
    d = oracle.diagnose(s)
    scale = np.abs(s.alpha).sum()
    assert np.abs(d[O_]).max() < 1e-15 * scale
    np.testing.assert_allclose(d[I_], [0, 0, math.pi * R * R * G], rtol=0, atol=1e-14)
    assert np.abs(d[A_]).max() < 1e-14
    c = 2.5
    t = s.copy()
    t.pos[:, 0] += c
    d = oracle.diagnose(t)
    np.testing.assert_allclose(d[A_], [0, -math.pi * G * R * R * c, 0], rtol=0, atol=1e-13)
    np.testing.assert_allclose(d[I_], [0, 0, math.pi * R * R * G], rtol=0, atol=1e-13)


def test_gaussian_enstrophy_is_zero(oracle):
    d = oracle.diagnose(W.ring_1k())
    assert d[E] == 0.0 and d[EF] == 0.0


def test_kinetic_energy_merges_coincident_particles(oracle):
    """Two coincident particles of strength a and core sigma carry the energy
    of one particle of strength 2a (the pair term at rho = 0 equals the self
    term): KE = (2a)^2 / (8 pi sigma), KEF = 3 (2a)^2 / (16 pi sigma)."""
    a = np.array([0.3, -0.2, 0.5])
    sig = 0.25
    s = W.ParticleSystem(off=np.array([0, 2], dtype=np.int64), pos=np.zeros((2, 3)),
                         alpha=np.stack([a, a]), radius=np.full(2, sig), volume=np.ones(2))
    s.pos[:] = [1.0, 2.0, 3.0]
    d = oracle.diagnose(s)
    a2 = float(4 * a @ a)
    assert d[KE] == pytest.approx(a2 / (8 * math.pi * sig), rel=1e-14)
    assert d[KEF] == pytest.approx(3 * a2 / (16 * math.pi * sig), rel=1e-14)


def test_kinetic_energy_far_field(oracle):
    """Two particles far apart: KE -> self terms + [a1.a2 + (x.a1)(x.a2)/rho^2]
    / (8 pi rho), the interaction energy of two singular vortex elements."""
    a1, a2 = np.array([0.3, 0.1, -0.4]), np.array([-0.2, 0.5, 0.1])
    x = np.array([30.0, -40.0, 120.0])
    sig = 0.1
    s = W.ParticleSystem(off=np.array([0, 2], dtype=np.int64), pos=np.stack([np.zeros(3), x]),
                         alpha=np.stack([a1, a2]), radius=np.full(2, sig), volume=np.ones(2))
    d = oracle.diagnose(s)
    rho = np.linalg.norm(x)
    self_ = (a1 @ a1 + a2 @ a2) / (8 * math.pi * sig)
    pair = (a1 @ a2 + (x @ a1) * (x @ a2) / rho ** 2) / (8 * math.pi * rho)
    assert d[KE] - self_ == pytest.approx(pair, rel=20 * (sig / rho) ** 2)  # O(sigma^2/rho^2) corrections


def test_helicity_is_minus_alpha_dot_velocity(oracle):
    """H = 2 sum_{i<j} q x_ij.(a_i x a_j) = -sum_i a_i . u_i with u the
    oracle's own (Biot-Savart-pinned) velocity."""
    rng = np.random.default_rng(2)
    n = 50
    s = W.ParticleSystem(off=np.array([0, 20, n], dtype=np.int64), pos=rng.normal(size=(n, 3)),
                         alpha=rng.normal(size=(n, 3)), radius=rng.uniform(0.2, 0.5, n),
                         volume=np.ones(n))
    u, _ = oracle.solve(s)
    d = oracle.diagnose(s)
    hel = -float(np.sum(s.alpha * u))
    assert d[H] == pytest.approx(hel, rel=1e-12, abs=1e-14 * np.sum(np.abs(s.alpha * u)))


def test_centroid(oracle):
    s = W.ring_1k()
    t = s.copy()
    t.pos[:, 2] += 0.75
    d = oracle.diagnose(t)
    # cross-section is symmetric in z about the ring plane
    assert d[ZC] == pytest.approx(0.75, rel=1e-14)
    # multi-wake: the last wake with a non-zero denominator wins
    lf = W.leapfrog_64k(nst=32)
    d = oracle.diagnose(lf)
    assert d[ZC] == pytest.approx(0.5, rel=1e-13)
    # no vorticity anywhere: keeps the previous value
    z = lf.copy()
    z.alpha[:] = 0.0
    assert oracle.diagnose(z, zc_prev=0.123)[ZC] == 0.123


def test_quadratic_diagnostics_do_not_depend_on_wake_order(oracle):
    lf = W.leapfrog_64k(nst=24)
    d1 = oracle.diagnose(lf)
    sw = W.concat([lf.wake(1), lf.wake(0)], coupled=True, nu=lf.nu, name="swapped")
    d2 = oracle.diagnose(sw)
    for m in (H, KE, KEF):
        assert d2[m] == pytest.approx(d1[m], rel=1e-13, abs=1e-15)
    np.testing.assert_allclose(d2[I_], d1[I_], rtol=1e-13, atol=1e-15)
