"""Pins of the oracle's vortex ring (vring.cpp:50-100, 223-242) and of its
strength resolution (__resolve::rebuild, resolve.cpp:12-63): geometry and
volumes of the generator in closed form, the influence matrix against a
direct numpy evaluation of the smoothed vorticity field, the pivoted LU
solve against LAPACK."""
import math

import numpy as np
import pytest

R, CORE, L, NAZI, SIG = 1.0, 0.1, 3, 49, 0.1924   # the reference's own ring (src/pawan.cpp:57)


def test_generator_geometry_and_volumes(oracle):
    pos, alpha, radius, volume = oracle.vring(R, CORE, L, NAZI, SIG)
    m = (2 * L + 1) ** 2
    assert len(pos) == NAZI * m and np.all(alpha == 0.0) and np.all(radius == SIG)
    r0 = 0.35 / (2 * L + 1)
    dpsi = 2 * math.pi / NAZI
    for a in (0, 7, NAZI - 1):
        st = pos[a * m:(a + 1) * m]
        psi = a * dpsi
        centre = np.array([R * math.cos(psi), R * math.sin(psi), 0.0])
        np.testing.assert_allclose(st[0], centre, atol=1e-15)
        # layer j: 8j particles on a circle of radius r_j about the core centre
        q = 1
        for j in range(1, L + 1):
            rj = r0 * (12 * j * j + 1) / (6 * j)
            d = np.linalg.norm(st[q:q + 8 * j] - centre, axis=1)
            np.testing.assert_allclose(d, rj, rtol=1e-14)
            q += 8 * j
        # the cell volumes telescope to the annulus: pi (0.35)^2 R dpsi
        assert volume[a * m:(a + 1) * m].sum() == pytest.approx(math.pi * 0.35 ** 2 * R * dpsi, rel=1e-13)


def test_ideal_profile(oracle):
    rate = oracle.vring_ideal(R, CORE, L)
    mag0 = 0.5 / (math.pi * CORE ** 2)
    assert rate[0] == pytest.approx(mag0, rel=1e-15)
    r0 = 0.35 / (2 * L + 1)
    r1 = r0 * 13 / 6
    # layer 1, theta = 0 and pi
    assert rate[1] == pytest.approx(mag0 * (1 + r1 / R) * math.exp(-0.5 * (r1 / CORE) ** 2), rel=1e-14)
    assert rate[5] == pytest.approx(mag0 * (1 - r1 / R) * math.exp(-0.5 * (r1 / CORE) ** 2), rel=1e-14)


def _field_numpy(pos, radius, m, nazi, p, i, zeta):
    """|(omega_x, omega_y)| at particle p when cell i of every station has
    strength 1: omega(x) = sum_q zeta_sigma(x - x_q) a_q."""
    q = np.arange(nazi) * m + i
    psi = np.arange(nazi) * 2 * np.pi / nazi
    t = np.stack([-np.sin(psi), np.cos(psi)], axis=1)
    rho = np.linalg.norm(pos[q] - pos[p], axis=1)
    sig = np.sqrt(0.5 * (radius[p] ** 2 + radius[q] ** 2))
    sig[q == p] = radius[p]
    f = (zeta(rho, sig)[:, None] * t).sum(axis=0)
    return math.hypot(f[0], f[1])


def _zeta_gauss(rho, s):
    return np.exp(-0.5 * (rho / s) ** 2) / (2 * np.pi) ** 1.5 / s ** 3


def _zeta_ho(rho, s):
    return 15.0 / (8 * np.pi) / ((rho / s) ** 2 + 1.0) ** 3.5 / s ** 3


@pytest.mark.parametrize("kernel", ["gaussian", "highorder"])
def test_influence_is_the_smoothed_field(oracle, kernel):
    nl, nazi = 2, 24
    pos, _, radius, _ = oracle.vring(R, CORE, nl, nazi, 0.12)
    m = (2 * nl + 1) ** 2
    kind, zeta = (oracle.GAUSSIAN, _zeta_gauss) if kernel == "gaussian" else (oracle.HIGHORDER, _zeta_ho)
    with oracle.smoothing(kind):
        M = oracle.vring_influence(nl, nazi, pos, radius)
    ref = np.array([[_field_numpy(pos, radius, m, nazi, p, i, zeta) for i in range(m)] for p in range(m)])
    np.testing.assert_allclose(M, ref, rtol=1e-13, atol=1e-13 * ref.max())


def test_lu_solve_matches_lapack(oracle):
    rng = np.random.default_rng(3)
    for n in (1, 5, 40, 97):
        A = rng.normal(size=(n, n)) + n * np.eye(n) * rng.uniform(0, 1)
        b = rng.normal(size=n)
        x = oracle.lu_solve(A, b)
        np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-10, atol=1e-12 * np.abs(x).max())
    # a pivot is needed: zero leading entry
    A = np.array([[0.0, 2.0], [3.0, 1.0]])
    np.testing.assert_allclose(oracle.lu_solve(A, [2.0, 4.0]), [1.0, 1.0], rtol=1e-15)
    with pytest.raises(ValueError):
        oracle.lu_solve(np.zeros((3, 3)), np.ones(3))


def test_lu_solve_flags_non_finite_pivots(oracle):
    """A column of NaN (what a zero smoothing radius gives: 0/0 in every
    entry) keeps the diagonal row as pivot and is reported as singular, as
    is an infinite pivot; a NaN below a finite pivot is never chosen."""
    rng = np.random.default_rng(4)
    A = rng.normal(size=(6, 6)) + 6 * np.eye(6)
    for col, val in ((0, np.nan), (3, np.nan), (2, np.inf)):
        B = A.copy()
        B[:, col] = val
        with pytest.raises(ValueError):
            oracle.lu_solve(B, np.ones(6))
    # the whole matrix NaN: every pivot stays on the diagonal
    with pytest.raises(ValueError):
        oracle.lu_solve(np.full((4, 4), np.nan), np.ones(4))
    # a single NaN below the diagonal loses every pivot search (fabs(NaN) >
    # best is false) until its row is the last candidate left, whose pivot
    # is then NaN: reported too
    B = A.copy()
    B[4, 0] = np.nan
    with pytest.raises(ValueError):
        oracle.lu_solve(B, np.ones(6))


def test_resolved_ring_reproduces_the_profile(oracle):
    s, M, (pos, alpha, radius, volume) = oracle.vring_resolve(R, CORE, L, NAZI, SIG)
    ideal = oracle.vring_ideal(R, CORE, L)
    assert np.abs(M @ s - ideal).max() <= 1e-10 * np.abs(ideal).max()
    np.testing.assert_allclose(s, np.linalg.solve(M, ideal), rtol=1e-6)  # cond(M) ~ 1.5e8
    m = len(s)
    np.testing.assert_allclose(np.linalg.norm(alpha[:m], axis=1), np.abs(s), rtol=1e-15)
