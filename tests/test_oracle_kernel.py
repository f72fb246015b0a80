"""Pins of the oracle's Gaussian kernel (gaussian.h:20-60) that do not
retype its formula: quadrature, moments, finite differences, limits and the
__float128 twin."""
import math

import numpy as np
import pytest
from scipy import integrate

K = (2.0 * math.pi) ** -1.5


def test_dnrm2_matches_hypot(oracle):
    rng = np.random.default_rng(1)
    for _ in range(200):
        x = rng.normal(size=3) * 10 ** rng.uniform(-5, 5)
        ref = math.sqrt(math.fsum(float(v) ** 2 for v in x))
        assert abs(oracle.dnrm2(x) - ref) <= 2 * np.spacing(ref)
    assert oracle.dnrm2([3.0, 4.0, 12.0]) == 13.0
    assert oracle.dnrm2([0.0, 0.0, 0.0]) == 0.0
    # scaled algorithm: no overflow or underflow of the squares
    assert oracle.dnrm2([3e300, 4e300, 0.0]) == pytest.approx(5e300, rel=1e-15)
    assert oracle.dnrm2([3e-300, 4e-300, 0.0]) == pytest.approx(5e-300, rel=1e-15)


def test_zeta_is_a_normalised_gaussian(oracle):
    """Z integrates to 1 over space and has second moment 3 sigma^2."""
    for sigma in (0.02, 0.3, 1.7):
        Z = lambda r: oracle.kernel(r, sigma)[2] if r > 0 else K / sigma ** 3
        m0 = integrate.quad(lambda r: 4 * math.pi * r * r * Z(r), 0, 40 * sigma, epsabs=0, epsrel=1e-13,
                            limit=200)[0]
        m2 = integrate.quad(lambda r: 4 * math.pi * r ** 4 * Z(r), 0, 40 * sigma, epsabs=0, epsrel=1e-13,
                            limit=200)[0]
        assert m0 == pytest.approx(1.0, rel=1e-11)
        assert m2 == pytest.approx(3 * sigma ** 2, rel=1e-11)


def test_q_is_the_enclosed_vorticity_fraction(oracle):
    """q(rho) = (1/rho^3) int_0^rho Z(r) r^2 dr (Gauss's law for the smoothed
    Biot-Savart kernel: q 4 pi rho^3 is the vorticity enclosed by rho)."""
    for sigma in (0.05, 1.0):
        for rb in (0.3, 0.9, 1.7, 3.2, 6.0):
            rho = rb * sigma
            enc = integrate.quad(lambda r: oracle.kernel(r, sigma)[2] * r * r if r > 0 else 0.0,
                                 0, rho, epsabs=0, epsrel=1e-13, limit=200)[0]
            q = oracle.kernel(rho, sigma)[0]
            tol = 1e-11 + 3.0 / rb ** 2 * 1e-15  # FP64 cancellation in q grows like 3/rho_bar^2
            assert q == pytest.approx(enc / rho ** 3, rel=tol)


def test_F_is_dq_drho_over_rho(oracle):
    for sigma in (0.08, 1.0):
        for rb in (0.5, 1.0, 2.0, 4.0):
            rho = rb * sigma
            h = 1e-4 * rho
            qp = oracle.kernel(rho + h, sigma)[0]
            qm = oracle.kernel(rho - h, sigma)[0]
            dq = (qp - qm) / (2 * h)
            F = oracle.kernel(rho, sigma)[1]
            assert F == pytest.approx(dq / rho, rel=5e-7)


def test_kernel_limits(oracle):
    sigma = 0.3
    C = K / sigma ** 3
    # rho_bar -> 0: q -> C/3, F -> -C/(5 sigma^2), Z -> C (series A.2)
    rb = 1e-2
    q, F, Z, n = oracle.kernel(rb * sigma, sigma)
    assert q == pytest.approx(C * (1 / 3 - rb ** 2 / 10), rel=1e-9)
    assert F == pytest.approx(C / sigma ** 2 * (-1 / 5 + rb ** 2 / 14), rel=1e-5)
    assert Z == pytest.approx(C, rel=1e-4) and n == Z
    # rho_bar -> inf: singular Biot-Savart kernel
    rho = 12 * sigma
    q, F, Z, n = oracle.kernel(rho, sigma)
    assert q == pytest.approx(1 / (4 * math.pi * rho ** 3), rel=1e-14)
    assert F == pytest.approx(-3 / (4 * math.pi * rho ** 5), rel=1e-12)
    assert Z < 1e-30


def test_kernel_against_float128_twin(oracle):
    """FP64 kernel vs the __float128 evaluation of the same formulas: at
    moderate rho_bar to rounding, and at small rho_bar within the
    cancellation bound (error ~ 3/rho_bar^2 ulp for q, 15/rho_bar^4 for F;
    Z carries the rounding of its exponent, ~ rho_bar^2 ulp)."""
    sigma = 0.2
    for rb in np.concatenate([np.linspace(0.5, 9.0, 30), [0.05, 0.1, 0.2]]):
        rho = rb * sigma
        q, F, Z, _ = oracle.kernel(rho, sigma)
        qq, FF, ZZ = oracle.kernel_q128(rho, sigma)
        eps = np.finfo(float).eps
        assert abs(Z - ZZ) <= (8 + 2 * rb ** 2) * eps * abs(ZZ)
        assert abs(q - qq) <= (16 + 8 * 3 / rb ** 2) * eps * abs(qq)
        assert abs(F - FF) <= (16 + 8 * 15 / rb ** 4) * eps * abs(FF)


def test_qsig_equals_kernel_q(oracle):
    for rho in (0.01, 0.2, 1.3):
        assert oracle.qsig(rho, 0.4) == oracle.kernel(rho, 0.4)[0]


def test_kernel_singular_at_zero_like_the_reference(oracle):
    """rho = 0 gives NaN in the reference (no guard, gaussian.h:53-59); the
    oracle keeps that behaviour."""
    q, F, Z, n = oracle.kernel(0.0, 0.1)
    assert math.isnan(q) and math.isnan(F)
