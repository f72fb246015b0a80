"""Pins of the oracle's steppers (integration.cpp:126-136, rk3.cpp:12-65,
rk4.cpp:12-77)."""
import numpy as np
import pytest

from conftest import normwise
from pawan_amd import workloads as W


@pytest.mark.parametrize("scheme,poly", [
    ("euler", lambda z: 1 + z),
    ("rk3", lambda z: 1 + z + z ** 2 / 2 + z ** 3 / 6),
    ("rk4", lambda z: 1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24),
])
def test_linear_system_amplification(oracle, scheme, poly):
    """A fake linear system y' = lambda y (the __system seam the integrators
    call through) is advanced by exactly the scheme's stability polynomial."""
    lam = np.array([-3.0, -0.5, 0.25, 1.0, -7.5, 0.0])
    y0 = np.array([1.0, -2.0, 0.5, 3.0, 1e-3, 4.0])
    dt = 0.05
    y = oracle.step_generic(scheme, dt, 1, y0, lambda y: lam * y)
    np.testing.assert_allclose(y, poly(lam * dt) * y0, rtol=4e-16 * 8, atol=0)
    y10 = oracle.step_generic(scheme, dt, 10, y0, lambda y: lam * y)
    np.testing.assert_allclose(y10, poly(lam * dt) ** 10 * y0, rtol=1e-14, atol=0)


@pytest.mark.parametrize("scheme,order", [("euler", 1), ("rk3", 3), ("rk4", 4)])
def test_observed_order_on_a_ring(oracle, scheme, order):
    # two coaxial rings: each deforms the other, so velocities change in time
    s = W.leapfrog_64k(nst=12)
    s.nu = 0.0
    T = 0.4
    ref = oracle.step(s, "rk4", T / 256, 256)
    errs = []
    for n in (16, 32):
        out = oracle.step(s, scheme, T / n, n)
        errs.append(normwise(out.pos - s.pos, ref.pos - s.pos))  # error of the displacement
    rate = np.log2(errs[0] / errs[1])
    assert rate == pytest.approx(order, abs=0.35)


def test_step_of_zero_dt_is_identity(oracle):
    s = W.ring_1k()
    out = oracle.step(s, "rk4", 0.0, 1)
    np.testing.assert_array_equal(out.pos, s.pos)
    np.testing.assert_array_equal(out.alpha, s.alpha)
