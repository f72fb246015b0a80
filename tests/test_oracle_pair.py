"""Pins of the oracle's pair interaction INTERACT
(interaction_utils.h:456-501): each term is checked against an independent
statement of the physics it computes."""
import math

import numpy as np
import pytest

from pawan_amd import workloads as W


def _pair(oracle, nu, ss, st, xs, xt, a_s, a_t, vs=1.0, vt=1.0):
    return oracle.interact(nu, ss, st, xs, xt, a_s, a_t, vs, vt)


def test_antisymmetric_vorticity_rate(oracle):
    rng = np.random.default_rng(3)
    for _ in range(50):
        xs, xt, a_s, a_t = (rng.normal(size=3) for _ in range(4))
        dr_s, dr_t, da_s, da_t = _pair(oracle, 0.01, 0.4, 0.7, xs, xt, a_s, a_t, 0.3, 1.3)
        np.testing.assert_array_equal(da_s, -da_t)


def test_far_field_velocity_is_biot_savart(oracle):
    """For rho >> sigma the velocity is the singular Biot-Savart law
    u(x) = a_s x (x - x_s) / (4 pi |x - x_s|^3)."""
    rng = np.random.default_rng(4)
    for _ in range(20):
        xs = rng.normal(size=3)
        xt = xs + 20.0 * rng.normal(size=3)
        a_s, a_t = rng.normal(size=3), rng.normal(size=3)
        dr_s, dr_t, _, _ = _pair(oracle, 0.0, 0.1, 0.1, xs, xt, a_s, a_t)
        d = xt - xs
        bs = np.cross(a_s, d) / (4 * math.pi * np.linalg.norm(d) ** 3)
        np.testing.assert_allclose(dr_t, bs, rtol=1e-12, atol=0)
        bs_s = np.cross(a_t, -d) / (4 * math.pi * np.linalg.norm(d) ** 3)
        np.testing.assert_allclose(dr_s, bs_s, rtol=1e-12, atol=0)


def test_stretching_is_transposed_velocity_gradient(oracle):
    """The transposed scheme: da_t/dt = (grad u)^T a_t, i.e.
    sum_i a_t,i du_i/dx_j, with u the velocity the source induces at the
    target.  Checked against a central finite difference of the oracle's own
    velocity (which the Biot-Savart test pins)."""
    rng = np.random.default_rng(5)
    for rb in (0.3, 1.0, 2.5, 8.0):
        sigma = 0.2
        xs = rng.normal(size=3)
        dirv = rng.normal(size=3)
        xt = xs + rb * sigma * dirv / np.linalg.norm(dirv)
        a_s, a_t = rng.normal(size=3), rng.normal(size=3)
        _, _, _, da_t = _pair(oracle, 0.0, sigma, sigma, xs, xt, a_s, a_t)
        h = 1e-5 * sigma
        grad_t = np.zeros(3)
        for j in range(3):
            e = np.zeros(3)
            e[j] = h
            up = _pair(oracle, 0.0, sigma, sigma, xs, xt + e, a_s, a_t)[1]
            um = _pair(oracle, 0.0, sigma, sigma, xs, xt - e, a_s, a_t)[1]
            grad_t[j] = a_t @ ((up - um) / (2 * h))
        np.testing.assert_allclose(da_t, grad_t, rtol=2e-6, atol=2e-6 * np.abs(grad_t).max())


def test_pse_is_viscous_laplacian_on_a_lattice(oracle):
    """Particle strength exchange: on a lattice of spacing h carrying
    a = omega(x) V with omega = (0, 0, cos(k x)), the vorticity rate of a
    particle is V nu (2/sigma^2) (exp(-k^2 sigma^2/2) - 1) cos(k x)
    (the Gaussian-weighted exchange integral in closed form; -> V nu
    laplacian(omega) as k sigma -> 0).  Parallel strengths carry no
    stretching, so the rate is the PSE term alone."""
    sigma, nu, k = 0.15, 0.7, 2.3
    h = sigma / 1.5
    m = 9
    g = np.arange(-m, m + 1) * h
    X, Y, Zc = np.meshgrid(g + 0.3, g, g, indexing="ij")
    pos = np.stack([X.ravel(), Y.ravel(), Zc.ravel()], axis=-1)
    V = h ** 3
    alpha = np.zeros_like(pos)
    alpha[:, 2] = np.cos(k * pos[:, 0]) * V
    n = len(pos)
    s = W.ParticleSystem(off=np.array([0, n], dtype=np.int64), pos=np.ascontiguousarray(pos),
                         alpha=alpha, radius=np.full(n, sigma), volume=np.full(n, V), nu=nu)
    centre = n // 2
    _, w = oracle.rates_targets(s, [centre])
    expect = V * nu * 2 / sigma ** 2 * (math.exp(-0.5 * (k * sigma) ** 2) - 1.0) * math.cos(k * pos[centre, 0])
    assert w[0, 0] == 0.0 and w[0, 1] == 0.0
    assert w[0, 2] == pytest.approx(expect, rel=1e-8)
    # and the continuum limit nu * laplacian(omega) * V to O((k sigma)^2)
    lap = -k * k * math.cos(k * pos[centre, 0]) * V * nu
    assert w[0, 2] == pytest.approx(lap, rel=(k * sigma) ** 2)


def test_rates_are_rotation_equivariant(oracle):
    rng = np.random.default_rng(6)
    s = W.ParticleSystem(off=np.array([0, 40], dtype=np.int64), pos=rng.normal(size=(40, 3)),
                         alpha=rng.normal(size=(40, 3)), radius=rng.uniform(0.2, 0.5, 40),
                         volume=rng.uniform(0.1, 1, 40), nu=0.05)
    u, w = oracle.solve(s)
    Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    r = s.copy()
    r.pos = np.ascontiguousarray(s.pos @ Q.T)
    r.alpha = np.ascontiguousarray(s.alpha @ Q.T)
    ur, wr = oracle.solve(r)
    np.testing.assert_allclose(ur, u @ Q.T, rtol=0, atol=1e-13 * np.abs(u).max())
    # a proper rotation leaves the pseudovector algebra unchanged; an
    # improper one would flip the sign of u (a x d) but not of da
    if np.linalg.det(Q) > 0:
        np.testing.assert_allclose(wr, w @ Q.T, rtol=0, atol=1e-13 * np.abs(w).max())


def test_rates_are_translation_invariant(oracle):
    rng = np.random.default_rng(7)
    s = W.ParticleSystem(off=np.array([0, 30], dtype=np.int64), pos=rng.normal(size=(30, 3)),
                         alpha=rng.normal(size=(30, 3)), radius=np.full(30, 0.3),
                         volume=np.full(30, 0.2), nu=0.01)
    u, w = oracle.solve(s)
    t = s.copy()
    t.pos = t.pos + np.array([0.5, -0.25, 0.125])  # exact in binary
    ut, wt = oracle.solve(t)
    np.testing.assert_allclose(ut, u, rtol=0, atol=1e-13 * np.abs(u).max())
    np.testing.assert_allclose(wt, w, rtol=0, atol=1e-13 * np.abs(w).max())


def test_linearity_in_source_strength(oracle):
    rng = np.random.default_rng(8)
    xs, xt, a_s, a_t = (rng.normal(size=3) for _ in range(4))
    r1 = _pair(oracle, 0.0, 0.3, 0.3, xs, xt, a_s, a_t)
    r2 = _pair(oracle, 0.0, 0.3, 0.3, xs, xt, 2 * a_s, a_t)
    np.testing.assert_array_equal(r2[1], 2 * r1[1])   # target velocity
    np.testing.assert_array_equal(r2[3], 2 * r1[3])   # stretching is bilinear
    # PSE is linear in nu
    p1 = _pair(oracle, 0.25, 0.3, 0.3, xs, xt, a_s, a_t, 0.5, 2.0)[3] - r1[3]
    p2 = _pair(oracle, 0.5, 0.3, 0.3, xs, xt, a_s, a_t, 0.5, 2.0)[3] - r1[3]
    np.testing.assert_allclose(p2, 2 * p1, rtol=1e-12, atol=4e-16 * np.abs(r1[3]).max())
    assert np.abs(p1).max() > 1e3 * 4e-16 * np.abs(r1[3]).max()
