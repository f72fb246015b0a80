"""Pins of the far-field oracle (oracle/farfield.py) by mathematics: the
expansion against exact cluster sums, its gradient against finite
differences, Morton codes of hand-placed points, and the network gradient
against finite differences."""
import math

import numpy as np
import pytest

from oracle import farfield as FF


def _cluster(rng, k, scale=1.0):
    pos = rng.normal(size=(k, 3)) * scale
    alpha = rng.normal(size=(k, 3))
    sigma = np.full(k, 1e-3)
    return pos, alpha, sigma


def _md_of(pos, alpha):
    c = pos.mean(axis=0)
    d = pos - c
    return c, alpha.sum(axis=0), alpha.T @ d


def test_md_single_particle_is_exact():
    """One particle at the centre: D = 0 and the monopole is the singular
    kernel itself (the Gaussian terms vanish far out)."""
    x = np.array([[3.0, -1.0, 2.0]])
    a = np.array([0.3, -0.7, 1.1])
    u, G = FF.md_eval(x, a, np.zeros((3, 3)))
    ue, Ge = FF.exact_node(x[0], 1e-3, np.zeros((1, 3)), a[None, :], np.array([1e-3]))
    np.testing.assert_allclose(u[0], ue, rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(G[0], Ge, rtol=1e-12, atol=1e-15)


def test_md_gradient_is_derivative_of_velocity():
    rng = np.random.default_rng(3)
    M = rng.normal(size=3)
    D = rng.normal(size=(3, 3))
    x = np.array([2.0, 1.5, -2.5])
    _, G = FF.md_eval(x[None], M, D)
    h = 1e-5
    fd = np.zeros((3, 3))
    for m in range(3):
        e = np.zeros(3)
        e[m] = h
        up, _ = FF.md_eval((x + e)[None], M, D)
        um, _ = FF.md_eval((x - e)[None], M, D)
        fd[:, m] = (up[0] - um[0]) / (2 * h)
    np.testing.assert_allclose(G[0], fd, rtol=1e-8, atol=1e-10)


@pytest.mark.parametrize("seed", [1, 2])
def test_md_error_is_second_order(seed):
    """Shrinking the cluster by 2 shrinks the expansion error by ~4 (the
    quadrupole is the first term left out); a missing or wrong dipole would
    leave an O(eps) error (ratio ~2)."""
    rng = np.random.default_rng(seed)
    pos0, alpha, sigma = _cluster(rng, 40)
    x = np.array([7.0, -5.0, 6.0])
    errs = []
    for eps in (0.2, 0.1, 0.05):
        pos = pos0 * eps
        c, M, D = _md_of(pos, alpha)
        u, G = FF.md_eval((x - c)[None], M, D)
        ue, Ge = FF.exact_node(x, 1e-3, pos, alpha, sigma)
        errs.append((np.linalg.norm(u[0] - ue), np.linalg.norm(G[0] - Ge)))
    for k in range(2):
        for j in range(2):
            ratio = errs[k][j] / errs[k + 1][j]
            assert 3.3 < ratio < 4.8, (k, j, ratio)


def test_md_without_dipole_is_first_order():
    """The same cluster evaluated with the dipole dropped: the error ratio
    is ~2, so the second-order test above does detect a missing term."""
    rng = np.random.default_rng(5)
    pos0, alpha, sigma = _cluster(rng, 40)
    x = np.array([7.0, -5.0, 6.0])
    errs = []
    for eps in (0.1, 0.05):
        pos = pos0 * eps
        c, M, _ = _md_of(pos, alpha)
        u, _ = FF.md_eval((x - c)[None], M, np.zeros((3, 3)))
        ue, _ = FF.exact_node(x, 1e-3, pos, alpha, sigma)
        errs.append(np.linalg.norm(u[0] - ue))
    assert 1.6 < errs[0] / errs[1] < 2.4


def _q_of(pos, alpha):
    c = pos.mean(axis=0)
    d = pos - c
    return np.einsum("sp,sj,sl->pjl", alpha, d, d)


@pytest.mark.parametrize("seed", [1, 2])
def test_quadrupole_error_is_third_order(seed):
    """With the quadrupole added, shrinking the cluster by 2 shrinks the
    error of velocity and gradient by ~8 (the octupole is the first term
    left out); the monopole + dipole alone fall by ~4 on the same clusters,
    so a dropped or wrong quadrupole term fails the bar."""
    rng = np.random.default_rng(seed)
    pos0, alpha, sigma = _cluster(rng, 40)
    x = np.array([7.0, -5.0, 6.0])
    errs = []
    for eps in (0.2, 0.1, 0.05):
        pos = pos0 * eps
        c, M, D = _md_of(pos, alpha)
        u, G = FF.md_eval((x - c)[None], M, D)
        uq, Gq = FF.quad_eval((x - c)[None], _q_of(pos, alpha))
        ue, Ge = FF.exact_node(x, 1e-3, pos, alpha, sigma)
        errs.append((np.linalg.norm(u[0] + uq[0] - ue), np.linalg.norm(G[0] + Gq[0] - Ge)))
    for k in range(2):
        for j in range(2):
            ratio = errs[k][j] / errs[k + 1][j]
            assert 6.5 < ratio < 9.5, (k, j, ratio)


def test_quadrupole_gradient_is_derivative_of_velocity():
    rng = np.random.default_rng(4)
    Q = rng.normal(size=(3, 3, 3))
    Q = 0.5 * (Q + Q.transpose(0, 2, 1))  # symmetric in the two offsets
    x = np.array([2.0, 1.5, -2.5])
    _, G = FF.quad_eval(x[None], Q)
    h = 1e-5
    fd = np.zeros((3, 3))
    for m in range(3):
        e = np.zeros(3)
        e[m] = h
        up, _ = FF.quad_eval((x + e)[None], Q)
        um, _ = FF.quad_eval((x - e)[None], Q)
        fd[:, m] = (up[0] - um[0]) / (2 * h)
    np.testing.assert_allclose(G[0], fd, rtol=1e-7, atol=1e-10)


def test_quadrupole_of_one_particle_is_half_its_second_difference():
    """One particle at offset delta from the centre: the quadrupole term is
    (1/2) (delta . grad)^2 of its field a x r/|r|^3, here by a central second
    difference along delta."""
    a = np.array([0.3, -0.7, 1.1])
    delta = np.array([0.02, -0.01, 0.015])
    x = np.array([3.0, -1.0, 2.0])
    uq, _ = FF.quad_eval(x[None], np.einsum("p,j,l->pjl", a, delta, delta))

    def field(y):
        return np.cross(a, y) / np.linalg.norm(y) ** 3

    h = 1.0
    sd = (field(x + h * delta) - 2.0 * field(x) + field(x - h * delta)) / (h * h)
    np.testing.assert_allclose(uq[0], 0.5 * sd, rtol=5e-4)


def test_transposed_contraction_matches_pair_stretching(oracle):
    """alpha_t^T G' / (4 pi) of a one-particle node equals the oracle's pair
    stretching (interaction_utils.h:456-501) for a far pair."""
    from pawan_amd.workloads import ParticleSystem
    xt, xs = np.array([0.0, 0.0, 0.0]), np.array([2.0, -1.0, 1.5])
    at, as_ = np.array([0.2, 0.5, -0.3]), np.array([-0.4, 0.1, 0.9])
    s = ParticleSystem(off=np.array([0, 2]), pos=np.array([xt, xs]), alpha=np.array([at, as_]),
                       radius=np.array([0.01, 0.01]), volume=np.array([1e-6, 1e-6]), nu=0.0,
                       coupled=True, name="pair")
    u, w = oracle.solve(s)
    ue, Ge = FF.exact_node(xt, 0.01, xs[None], as_[None], np.array([0.01]))
    np.testing.assert_allclose(ue / FF.FOUR_PI, u[0], rtol=1e-12)
    np.testing.assert_allclose(at @ Ge / FF.FOUR_PI, w[0], rtol=1e-11)


def test_morton_codes_of_hand_placed_points():
    pos = np.array([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
    codes = FF.morton_codes(pos)
    full = (1 << 30) - 1  # all 30 bits: 1022 on every axis is 0b1111111110
    q = 1022
    assert codes[0] == 0
    x = int(FF._spread10(np.array([q]))[0])
    assert codes[2] == x and codes[3] == x << 1 and codes[4] == x << 2
    assert codes[1] == x | (x << 1) | (x << 2)
    assert codes[1] <= full
    # interleaving: bit 3k of the code is bit k of the x cell index
    assert all(((x >> (3 * k)) & 1) == ((q >> k) & 1) for k in range(10))


def test_morton_perm_is_stable_on_ties():
    pos = np.array([[0.5, 0.5, 0.5]] * 4 + [[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]])
    perm = FF.morton_perm(pos)
    assert list(perm) == [4, 0, 1, 2, 3, 5]


def test_node_ranges_cover_levels():
    lf = np.array(list(range(0, 5000, 256)) + [5000])  # 20 leaves
    first, last, level, child = FF.node_ranges(lf)
    counts, offs = FF.tree_levels(len(lf) - 1)
    assert counts == [20, 3, 1]
    assert first[offs[1]] == 0 and last[offs[1]] == 2048
    assert first[offs[1] + 2] == 4096 and last[offs[1] + 2] == 5000
    assert last[-1] == 5000 and first[-1] == 0 and level[-1] == 2


def test_leaves_follow_octree_cells():
    """Codes of two far-apart clusters (600 and 300 particles, one cell each
    at level 1) and 20 singletons: every leaf lies in one cell, cells are cut
    into runs of 256."""
    a = np.full(600, 0b001 << 27, dtype=np.uint64) + np.arange(600, dtype=np.uint64)
    b = np.full(300, 0b110 << 27, dtype=np.uint64) + np.arange(300, dtype=np.uint64)
    c = (np.arange(20, dtype=np.uint64) + 2) << np.uint64(24)
    codes = np.sort(np.concatenate([a, b, c]))
    lf = FF.leaf_first(codes)
    assert lf[-1] == len(codes) and np.all(np.diff(lf) > 0) and np.all(np.diff(lf) <= 256)
    # no leaf mixes codes of different level-1 cells
    for x, y in zip(lf[:-1], lf[1:]):
        assert len(set((codes[x:y] >> np.uint64(27)).tolist())) == 1
    # level 1 is chosen (920 particles, 4 occupied cells); cell 1 holds the
    # 600-cluster and 8 of the singletons: 608 = 256 + 256 + 96; cell 6 holds
    # the 300-cluster: 256 + 44; cells 0 and 2 hold 6 singletons each
    np.testing.assert_array_equal(np.diff(lf), [6, 256, 256, 96, 6, 256, 44])


def test_prefix_levels():
    codes = np.array([0, 1, 8, 1 << 27, (1 << 27) + 1, (1 << 27) + 1], dtype=np.uint64)
    np.testing.assert_array_equal(FF.prefix_levels(codes), [9, 8, 0, 9, 10])


def _finite_diff(params, X, Y, idx, h=1e-6):
    out = []
    for i in idx:
        p = params.copy()
        p[i] += h
        lp, _ = FF.mlp_loss_grad(p, X, Y)
        p[i] -= 2 * h
        lm, _ = FF.mlp_loss_grad(p, X, Y)
        out.append((lp - lm) / (2 * h))
    return np.array(out)


def test_mlp_gradient_matches_finite_differences():
    rng = np.random.default_rng(11)
    params = rng.normal(size=FF.N_PARAMS) * 0.2
    X = rng.normal(size=(7, FF.N_IN))
    Y = rng.normal(size=(7, FF.N_OUT))
    loss, g = FF.mlp_loss_grad(params, X, Y)
    # a few entries of every parameter block
    idx = []
    for name, (o, size) in FF.OFF.items():
        idx += list(o + rng.choice(size, size=4, replace=False))
    fd = _finite_diff(params, X, Y, idx)
    np.testing.assert_allclose(g[idx], fd, rtol=2e-6, atol=1e-8)
    assert loss == pytest.approx(float(((FF.mlp_forward(params, X) - Y) ** 2).sum() / 7))


def test_gelu_derivative():
    x = np.linspace(-6, 6, 101)
    h = 1e-6
    np.testing.assert_allclose(FF.gelu_d(x), (FF.gelu(x + h) - FF.gelu(x - h)) / (2 * h), rtol=1e-7, atol=1e-10)
    # the tanh form 0.5 x (1 + tanh z)
    z = FF.C_GELU * (x + 0.044715 * x ** 3)
    np.testing.assert_allclose(FF.gelu(x), 0.5 * x * (1 + np.tanh(z)), rtol=1e-12, atol=1e-15)


def test_bf16_rounding():
    assert FF.bf16(np.array([1.0]))[0] == 1.0
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties to even (1.0)
    assert FF.bf16(np.array([1.0 + 2 ** -8]))[0] == 1.0
    assert FF.bf16(np.array([1.0 + 3 * 2 ** -8]))[0] == 1.0 + 2 ** -6
    assert FF.bf16(np.array([-2.0 ** -130]))[0] == -(2.0 ** -130)  # subnormal kept


def test_frame_is_a_rotation_fixed_by_moment_and_offset():
    rng = np.random.default_rng(2)
    for _ in range(20):
        r, M = rng.normal(size=3), rng.normal(size=3)
        R = FF.frame(r, M, 1.0)
        np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-12)
        assert np.linalg.det(R) == pytest.approx(1.0)
        np.testing.assert_allclose((R @ M)[1:], 0.0, atol=1e-12)  # M along e1
        assert abs((R @ r)[2]) < 1e-12 and (R @ r)[1] > 0      # r in the (e1, e2) half-plane
    # tiny moment: e1 along r; r parallel to M: any normal completes the frame
    R = FF.frame(np.array([0.0, 2.0, 0.0]), np.array([1e-9, 0, 0]), 1.0)
    np.testing.assert_allclose(R[0], [0, 1, 0], atol=1e-12)
    R = FF.frame(np.array([0.0, 0.0, 3.0]), np.array([0.0, 0.0, 1.0]), 1.0)
    np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-12)


def test_features_are_rotation_invariant():
    """Rotating the whole configuration leaves the network inputs unchanged
    and rotates nothing else: the frame absorbs the rotation."""
    rng = np.random.default_rng(4)
    r, M, D = rng.normal(size=(1, 3)), rng.normal(size=3), rng.normal(size=(3, 3))
    Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    if np.linalg.det(Q) < 0:
        Q[:, 0] *= -1
    X1, su1, sG1, _ = FF.features(r, 0.3, 2.0, M, D, 0.02)
    X2, su2, sG2, _ = FF.features(r @ Q.T, 0.3, 2.0, Q @ M, Q @ D @ Q.T, 0.02)
    np.testing.assert_allclose(X1, X2, atol=1e-12)
    np.testing.assert_allclose(su1, su2)
