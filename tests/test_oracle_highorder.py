"""Pins of the oracle's high-order algebraic smoothing kernel
(src/interaction/highorder.h:20-106), the compile-time alternative to the
Gaussian (interaction_utils.h:22-26).  Each function is checked against an
independent statement of what it must be: normalisation and moments of the
core, the velocity kernel as the integral of the core, F as q'/rho, the PSE
kernel by the exchange integral in Fourier space, and the pair and
diagnostic identities that hold for any smoothing kernel."""
import math

import numpy as np
import pytest
from scipy import integrate

from pawan_amd import workloads as W

SIG = 0.7


@pytest.fixture
def ho(oracle):
    with oracle.smoothing(oracle.HIGHORDER):
        yield oracle


def test_core_is_normalised(ho):
    tot, _ = integrate.quad(lambda r: 4 * math.pi * r * r * ho.kernel(r, SIG)[2], 0, np.inf,
                            epsabs=0, epsrel=1e-12, limit=200)
    assert tot == pytest.approx(1.0, rel=1e-10)


def test_q_is_the_enclosed_core(ho):
    """q(rho) = (1/rho^3) int_0^rho Z(r) r^2 dr: the velocity kernel of a
    smoothed vortex is the singular one times the enclosed vorticity."""
    for rb in (0.05, 0.3, 1.0, 2.2, 6.0):
        rho = rb * SIG
        enc, _ = integrate.quad(lambda r: ho.kernel(r, SIG)[2] * r * r, 0, rho, epsabs=0, epsrel=1e-13)
        assert ho.kernel(rho, SIG)[0] == pytest.approx(enc / rho ** 3, rel=1e-11)
        assert ho.qsig(rho, SIG) == ho.kernel(rho, SIG)[0]


def test_f_is_the_radial_derivative_of_q(ho):
    for rb in (0.2, 0.8, 1.5, 4.0):
        rho = rb * SIG
        h = 1e-5 * rho
        dq = (ho.kernel(rho + h, SIG)[0] - ho.kernel(rho - h, SIG)[0]) / (2 * h)
        assert ho.kernel(rho, SIG)[1] == pytest.approx(dq / rho, rel=1e-8)


def test_far_velocity_error_is_fourth_order(ho):
    """q 4 pi rho^3 = 1 - (15/8) (sigma/rho)^4 + O((sigma/rho)^6): the
    kernel's velocity differs from Biot-Savart at fourth order, the property
    that names it (the Gaussian's error is exponentially small instead)."""
    for rb in (20.0, 40.0):
        rho = rb * SIG
        err = 1.0 - 4 * math.pi * rho ** 3 * ho.kernel(rho, SIG)[0]
        assert err * rb ** 4 == pytest.approx(1.875, rel=5.0 / rb ** 2)


def _eta_hat(oracle, k, sigma):
    """Fourier transform of the PSE kernel eta_sigma at wavenumber k."""
    f = lambda r: 4 * math.pi * r * r * oracle.kernel(r, sigma)[3] * (math.sin(k * r) / (k * r) if r > 0 else 1.0)
    v, _ = integrate.quad(f, 0, 60 * sigma, epsabs=0, epsrel=1e-12, limit=400)
    return v


def test_pse_kernel_second_moment_gives_the_laplacian(ho):
    """int eta rho_bar^2 dV = 3 makes (2/sigma^2) int (a(y) - a(x)) eta the
    Laplacian (the PSE consistency condition)."""
    m2, _ = integrate.quad(lambda r: 4 * math.pi * r ** 4 / SIG ** 2 * ho.kernel(r, SIG)[3], 0, np.inf,
                           epsabs=0, epsrel=1e-12, limit=200)
    assert m2 == pytest.approx(3.0, rel=1e-10)


def test_pse_on_a_lattice_is_the_exchange_integral(ho):
    """On a lattice carrying a = omega(x) V with omega = (0, 0, cos(k x)), the
    rate of a particle is V nu (2/sigma^2) (eta_hat(k) - eta_hat(0)) cos(k x),
    eta_hat by quadrature; and -> V nu laplacian(omega) as k sigma -> 0."""
    # spacing sigma/3: the algebraic kernel's transform decays only as
    # exp(-k sigma), so the lattice needs a finer spacing than a Gaussian
    sigma, nu, k = 0.15, 0.7, 2.3
    h = sigma / 3.0
    m = 28
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
    _, w = ho.rates_targets(s, [centre])
    eta0 = _eta_hat(ho, 1e-9, sigma)
    expect = V * nu * 2 / sigma ** 2 * (_eta_hat(ho, k, sigma) - eta0) * math.cos(k * pos[centre, 0])
    assert w[0, 0] == 0.0 and w[0, 1] == 0.0
    # aliasing of the sigma/3 lattice and its truncation at 9 sigma
    assert w[0, 2] == pytest.approx(expect, rel=1e-3)
    lap = -k * k * math.cos(k * pos[centre, 0]) * V * nu
    assert w[0, 2] == pytest.approx(lap, rel=(k * sigma) ** 2)


def test_stretching_is_transposed_velocity_gradient(ho):
    rng = np.random.default_rng(15)
    for rb in (0.3, 1.0, 2.5, 8.0):
        sigma = 0.2
        xs = rng.normal(size=3)
        dirv = rng.normal(size=3)
        xt = xs + rb * sigma * dirv / np.linalg.norm(dirv)
        a_s, a_t = rng.normal(size=3), rng.normal(size=3)
        _, _, _, da_t = ho.interact(0.0, sigma, sigma, xs, xt, a_s, a_t, 1.0, 1.0)
        hh = 1e-5 * sigma
        grad_t = np.zeros(3)
        for j in range(3):
            e = np.zeros(3)
            e[j] = hh
            up = ho.interact(0.0, sigma, sigma, xs, xt + e, a_s, a_t, 1.0, 1.0)[1]
            um = ho.interact(0.0, sigma, sigma, xs, xt - e, a_s, a_t, 1.0, 1.0)[1]
            grad_t[j] = a_t @ ((up - um) / (2 * hh))
        np.testing.assert_allclose(da_t, grad_t, rtol=2e-6, atol=2e-6 * np.abs(grad_t).max())


def test_antisymmetric_vorticity_rate(ho):
    rng = np.random.default_rng(13)
    for _ in range(20):
        xs, xt, a_s, a_t = (rng.normal(size=3) for _ in range(4))
        _, _, da_s, da_t = ho.interact(0.01, 0.4, 0.7, xs, xt, a_s, a_t, 0.3, 1.3)
        np.testing.assert_array_equal(da_s, -da_t)


def test_enstrophy_factors(ho):
    """ENSTF's pair factor is the core Z itself; both pair factors meet the
    self terms as rho -> 0."""
    for rb in (0.0, 0.4, 1.3, 5.0):
        rho = rb * SIG
        if rho > 0:
            assert ho.enstf_pair(rho, SIG) == pytest.approx(ho.kernel(rho, SIG)[2], rel=1e-14)
    F1, F2 = ho.enst_pair(0.0, SIG)
    assert F1 == pytest.approx(ho.enst_self(SIG), rel=1e-14)
    assert F1 == pytest.approx(5.0 / (2.0 * math.pi * SIG ** 3), rel=1e-14)
    assert ho.enstf_pair(0.0, SIG) == pytest.approx(ho.enstf_self(SIG), rel=1e-14)


def test_enstrophy_pair_is_the_hessian_form_of_the_core_potential(ho):
    """ENST's pair term F1 a1.a2 + F2 (a1.x)(a2.x) is a1.(Z I - grad grad G) a2
    with G the Newton potential of the core (-lap G = Z):
        G(r) = (1/r) int_0^r Z s^2 ds + int_r^inf Z s ds,
        G' = -M/r^2 with M = int_0^r Z s^2 ds,  G'' = -Z - 2 G'/r,
    so F1 = Z - G'/r = Z + M/r^3 and F2 = -(G'' - G'/r)/r^2 = (Z - 3 M/r^3)/r^2.
    M and the outer integral come from quadrature of the core alone; the
    Laplacian of G is checked by finite differences of the quadrature, so a
    wrong coefficient in either factor's rho-dependence fails."""
    Z = lambda r: ho.kernel(r, SIG)[2]

    def potential(r):
        inner, _ = integrate.quad(lambda s: Z(s) * s * s, 0, r, epsabs=0, epsrel=1e-13)
        outer, _ = integrate.quad(lambda s: Z(s) * s, r, np.inf, epsabs=0, epsrel=1e-13, limit=200)
        return inner / r + outer, inner

    for rb in (0.15, 0.6, 1.0, 1.7, 3.0, 8.0):
        r = rb * SIG
        G, M = potential(r)
        h = 1e-2 * r
        g = [potential(r + k * h)[0] if k else G for k in (-2, -1, 0, 1, 2)]
        d1 = (g[0] - 8 * g[1] + 8 * g[3] - g[4]) / (12 * h)  # five-point stencils
        d2 = (-g[0] + 16 * g[1] - 30 * g[2] + 16 * g[3] - g[4]) / (12 * h * h)
        assert d1 == pytest.approx(-M / r ** 2, rel=1e-7)
        if rb <= 3.0:  # beyond, -lap G is a small difference of large terms
            assert -(d2 + 2 * d1 / r) == pytest.approx(Z(r), rel=1e-5)  # -lap G = Z
        F1, F2 = ho.enst_pair(r, SIG)
        assert F1 == pytest.approx(Z(r) + M / r ** 3, rel=1e-11)
        assert F2 == pytest.approx((Z(r) - 3 * M / r ** 3) / r ** 2, rel=1e-10)
    # far from the core: F1 -> 1/(4 pi r^3) and F2 r^2 -> -3/(4 pi r^3), the
    # singular Hessian of 1/(4 pi r)
    r = 400 * SIG
    F1, F2 = ho.enst_pair(r, SIG)
    assert F1 * 4 * math.pi * r ** 3 == pytest.approx(1.0, rel=1e-4)
    assert F2 * 4 * math.pi * r ** 5 == pytest.approx(-3.0, rel=1e-4)


def test_diagnostics(ho, oracle):
    """Helicity = -sum a.u with the oracle's own high-order velocity; the
    enstrophy sums are the self terms plus twice the pair terms of i < j."""
    rng = np.random.default_rng(12)
    n = 40
    s = W.ParticleSystem(off=np.array([0, 15, n], dtype=np.int64), pos=rng.normal(size=(n, 3)),
                         alpha=rng.normal(size=(n, 3)), radius=rng.uniform(0.2, 0.5, n),
                         volume=np.ones(n))
    u, _ = ho.solve(s)
    d = ho.diagnose(s)
    assert d[9] == pytest.approx(-float(np.sum(s.alpha * u)), rel=1e-12)
    E = EF = 0.0
    for i in range(n):
        a2 = s.alpha[i] @ s.alpha[i]
        E += ho.enst_self(s.radius[i]) * a2
        EF += ho.enstf_self(s.radius[i]) * a2
        for j in range(i + 1, n):
            x = s.pos[j] - s.pos[i]
            sg = math.sqrt(0.5 * (s.radius[i] ** 2 + s.radius[j] ** 2))
            F1, F2 = ho.enst_pair(np.linalg.norm(x), sg)
            E += 2 * (F1 * (s.alpha[i] @ s.alpha[j]) + F2 * (s.alpha[i] @ x) * (s.alpha[j] @ x))
            EF += 2 * ho.enstf_pair(np.linalg.norm(x), sg) * (s.alpha[i] @ s.alpha[j])
    assert d[10] == pytest.approx(E, rel=1e-12)
    assert d[11] == pytest.approx(EF, rel=1e-12)
    with oracle.smoothing(oracle.GAUSSIAN):
        dg = oracle.diagnose(s)
    assert dg[10] == 0.0 and dg[11] == 0.0
    assert dg[12] == d[12]   # kinetic energy does not depend on the kernel
