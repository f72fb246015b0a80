"""The oracle's __wake::split and __wake::merge (src/wake/wake.cpp:61-159)
pinned by hand-computed cases, the reference's loop quirks and invariants:
split conserves the total strength exactly and the linear impulse to
rounding, merge conserves the total strength to rounding; ages end the outer
loops; merge takes |a_i| once per outer particle."""
import numpy as np
import pytest

from oracle import oracle as O


def _wake(pos, alpha, radius, volume=None, birth=None, active=None, maxn=None):
    n = len(pos)
    return O.WakeParticles(np.asarray(pos, float), np.asarray(alpha, float), np.asarray(radius, float),
                           np.ones(n) if volume is None else np.asarray(volume, float),
                           None if birth is None else np.asarray(birth, float),
                           np.zeros(n) if active is None else np.asarray(active, float), maxn)


def test_split_one_particle_by_hand():
    # |a| = 5 >= 2 * birth = 4, birth > 1e-4, age 50 > 45
    w = _wake([[1.0, 2.0, 3.0]], [[3.0, 4.0, 0.0]], [0.1], volume=[0.7], birth=[2.0], active=[7.0], maxn=3)
    assert O.wake_split(w, 57) == 1
    assert (w.n, w.npidx) == (2, 2)
    d = np.array([3.0, 4.0, 0.0]) * (0.25 * 0.1 / 5.0)
    np.testing.assert_array_equal(w.pos[0], np.array([1.0, 2.0, 3.0]) - d)
    np.testing.assert_array_equal(w.pos[1], np.array([1.0, 2.0, 3.0]) + d)
    np.testing.assert_array_equal(w.alpha[:2], [[1.5, 2.0, 0.0], [1.5, 2.0, 0.0]])
    np.testing.assert_array_equal(w.birth[:2], [2.5, 2.5])
    np.testing.assert_array_equal(w.radius[:2], [0.1, 0.1])
    np.testing.assert_array_equal(w.volume[:2], [0.7, 0.7])
    np.testing.assert_array_equal(w.active[:2], [7.0, 7.0])


def test_split_tests_and_the_age_prefix():
    pos = np.arange(12, dtype=float).reshape(4, 3)
    alpha = [[1.0, 0, 0], [4.0, 0, 0], [1.0, 0, 0], [9.0, 0, 0]]
    # 0: |a| < 2 birth; 1: splits; 2: birth too small; 3: splits
    w = _wake(pos, alpha, [0.1] * 4, birth=[1.0, 1.0, 0.5e-4, 2.0], active=[0, 0, 0, 0], maxn=8)
    assert O.wake_split(w, 46) == 2 and w.n == 6
    np.testing.assert_array_equal(w.alpha[4], [2.0, 0, 0])  # parents' order
    np.testing.assert_array_equal(w.alpha[5], [4.5, 0, 0])
    # a particle of age 45 or less ends the loop: nothing after it splits
    w = _wake(pos, alpha, [0.1] * 4, birth=[1.0, 1.0, 0.5e-4, 2.0], active=[0, 5, 0, 0], maxn=8)
    assert O.wake_split(w, 50) == 0 and w.n == 4
    # a new particle is as old as its parent and never splits again
    w = _wake(pos, alpha, [0.1] * 4, birth=[1.0, 1.0, 0.5e-4, 2.0], active=[0, 0, 0, 0], maxn=8)
    O.wake_split(w, 46)
    assert O.wake_split(w, 47) == 0


def test_split_capacity_and_npidx():
    w = _wake([[0, 0, 0], [1, 0, 0]], [[5.0, 0, 0], [5.0, 0, 0]], [0.1, 0.1], birth=[1.0, 1.0], maxn=3)
    before = w.copy()
    assert O.wake_split(w, 100) == -1  # two new particles, one free row
    np.testing.assert_array_equal(w.pos, before.pos)
    w = _wake([[0, 0, 0]], [[5.0, 0, 0]], [0.1], birth=[1.0], maxn=4)
    w.npidx = 0  # the reference would overwrite particle 0
    assert O.wake_split(w, 100) == -1


def test_split_invariants_on_a_random_wake():
    rng = np.random.default_rng(4)
    n = 500
    pos = rng.normal(size=(n, 3))
    alpha = rng.normal(size=(n, 3))
    birth = np.linalg.norm(alpha, axis=1) * rng.uniform(0.2, 0.8, n)
    w = _wake(pos, alpha, rng.uniform(0.05, 0.2, n), birth=birth, active=np.zeros(n), maxn=2 * n)
    s0 = w.alpha[:n].sum(axis=0)
    i0 = np.cross(w.pos[:n], w.alpha[:n]).sum(axis=0)
    k = O.wake_split(w, 60)
    assert k == np.count_nonzero(np.linalg.norm(alpha, axis=1) >= 2 * birth) and w.n == n + k
    assert np.allclose(w.alpha[:w.n].sum(axis=0), s0, rtol=0, atol=1e-13)
    assert np.allclose(np.cross(w.pos[:w.n], w.alpha[:w.n]).sum(axis=0), i0, rtol=0, atol=1e-12)
    # every new particle is 0.5 sigma from its parent, along a
    par = np.flatnonzero(np.linalg.norm(alpha, axis=1) >= 2 * birth)
    sep = w.pos[n:w.n] - w.pos[par]
    np.testing.assert_allclose(np.linalg.norm(sep, axis=1), 0.5 * w.radius[par], rtol=1e-13)
    np.testing.assert_allclose(np.cross(sep, alpha[par]), 0.0, atol=1e-13)


def test_merge_one_pair_by_hand():
    # sigma 0.125 each: reach (0.125 + 0.125) / 8 = 0.03125 (exact in binary)
    w = _wake([[0, 0, 0], [0.03, 0, 0]], [[1.0, 0, 0], [0, 3.0, 0]], [0.125, 0.125], volume=[2.0, 5.0])
    assert O.wake_merge(w, 46) == 1
    np.testing.assert_array_equal(w.pos[0], [(0 * 1 + 0.03 * 3) * (1 / 4.0), 0, 0])
    np.testing.assert_array_equal(w.alpha[:2], [[1.0, 3.0, 0], [0, 0, 0]])
    np.testing.assert_array_equal(w.volume[:2], [2.0, 0.0])
    np.testing.assert_array_equal(w.pos[1], [0.03, 0, 0])  # kept in place


@pytest.mark.parametrize("dx,a1,merged", [
    (0.03125, [0, 1.0, 0], False),    # at the reach: not strictly inside
    (0.0312, [0, 1.0, 0], True),
    (0.01, [2.0, 0, 0], False),       # parallel: 1 - cos = 0 < 1e-4
    (0.01, [-2.0, 0, 0], True),       # antiparallel: 2
    (0.01, [1.0, 0.01, 0], False),    # 1 - cos = 5.0e-5
    (0.01, [1.0, 0.02, 0], True),     # 1 - cos = 2.0e-4
    (0.01, [0, 0, 0], False),         # zero strength
])
def test_merge_tests(dx, a1, merged):
    w = _wake([[0, 0, 0], [dx, 0, 0]], [[1.0, 0, 0], a1], [0.125, 0.125])
    assert O.wake_merge(w, 100) == (1 if merged else 0)


def test_merge_ages():
    pos = [[0, 0, 0], [0.01, 0, 0], [0.5, 0, 0], [0.51, 0, 0]]
    alpha = [[1.0, 0, 0], [0, 1.0, 0], [1.0, 0, 0], [0, 1.0, 0]]
    # particle 0 too young: the outer loop never starts
    w = _wake(pos, alpha, [0.125] * 4, active=[10, 0, 0, 0])
    assert O.wake_merge(w, 50) == 0
    # particle 1 young: 0 still absorbs it (the inner loop has no age test),
    # then the outer loop ends at 1 and 2, 3 stay apart
    w = _wake(pos, alpha, [0.125] * 4, active=[0, 10, 0, 0])
    assert O.wake_merge(w, 50) == 1
    np.testing.assert_array_equal(w.alpha[2:4], [[1.0, 0, 0], [0, 1.0, 0]])


def test_merge_chain_follows_the_moving_particle():
    # 0 absorbs 1 (|a_1| = 9 pulls x_0 to 0.018), which brings 2 within reach
    # (0.027 < 0.03125); the second merge weighs x_0 with the |a_0| taken
    # before the first (1, not |a_0 + a_1|)
    pos = [[0, 0, 0], [0.02, 0, 0], [0.045, 0, 0]]
    alpha = [[1.0, 0, 0], [0, 9.0, 0], [0, 0, 2.0]]
    w = _wake(pos, alpha, [0.125] * 3)
    assert O.wake_merge(w, 46) == 2
    x0 = (0 * 1.0 + 0.02 * 9.0) * (1 / 10.0)
    x0 = (x0 * 1.0 + 0.045 * 2.0) * (1 / 3.0)
    np.testing.assert_array_equal(w.pos[0], [x0, 0, 0])
    np.testing.assert_array_equal(w.alpha[0], [1.0, 9.0, 2.0])
    # without 1, 2 is out of reach of the unmoved 0
    w = _wake([pos[0], pos[2]], [alpha[0], alpha[2]], [0.125] * 2)
    assert O.wake_merge(w, 46) == 0


def test_merge_misalignment_uses_the_stale_norm():
    # a_0 grows to (1, 1, 0) by the first merge; the test against particle 2
    # takes 1 - a_0 . a_2 / |a_0 before| / |a_2| = 1 - 1.4 / 1 / 1 < 1e-4: no
    # merge (with the grown |a_0| = sqrt 2 it would be 0.01 >= 1e-4)
    pos = [[0, 0, 0], [0.01, 0, 0], [0.0, 0.01, 0]]
    alpha = [[1.0, 0, 0], [0, 1.0, 0], [0.6, 0.8, 0]]
    w = _wake(pos, alpha, [0.125] * 3)
    assert O.wake_merge(w, 46) == 1
    np.testing.assert_array_equal(w.alpha[2], [0.6, 0.8, 0])


def test_merge_conserves_the_total_strength():
    rng = np.random.default_rng(8)
    n = 400
    base = rng.uniform(-1, 1, size=(n // 4, 3))
    pos = np.repeat(base, 4, axis=0) + rng.normal(scale=0.004, size=(n, 3))  # clusters of 4
    alpha = rng.normal(size=(n, 3))
    w = _wake(pos, alpha, rng.uniform(0.05, 0.15, n), active=np.zeros(n))
    s0 = w.alpha[:n].sum(axis=0)
    k = O.wake_merge(w, 100)
    assert k > 50
    assert np.allclose(w.alpha[:n].sum(axis=0), s0, rtol=0, atol=1e-12)
    zero = np.flatnonzero(np.all(w.alpha[:n] == 0, axis=1))
    assert len(zero) == k and np.all(w.volume[zero] == 0)
