"""Particle management on the GPU (pw_wake_split, pw_wake_merge: __wake::split
and merge, src/wake/wake.cpp:61-159) against the oracle's sequential
restatement, bit for bit: wakes with planted clusters (merge chains), a
spread of radii, zero strengths, an age cut, then the 1M helical wake with
disturbed strengths (conservation of the total strength)."""
import math
import os
import subprocess
import time

import numpy as np
import pytest

from oracle import oracle as O
from pawan_amd import capi, workloads as W

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "bin", "relax_driver")


def _say(msg):
    print(msg, flush=True)


def _clustered(n, seed, spread=0.01):
    """n particles in clusters of 1-6 about random centres, strengths of
    random directions (some zero, some aligned pairs), radii 0.05-0.15,
    born at steps 0-10."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(1, 7, size=n)
    sizes = sizes[np.cumsum(sizes) <= n]
    sizes = np.append(sizes, n - sizes.sum()) if sizes.sum() < n else sizes
    centres = rng.uniform(-1.0, 1.0, size=(len(sizes), 3))
    pos = np.repeat(centres, sizes, axis=0) + rng.normal(scale=spread, size=(n, 3))
    alpha = rng.normal(size=(n, 3))
    alpha[rng.random(n) < 0.05] = 0.0
    al = rng.random(n) < 0.2  # some strengths aligned with the previous particle's
    al[0] = False
    alpha[al] = alpha[np.flatnonzero(al) - 1] * rng.uniform(0.5, 2.0, size=(al.sum(), 1))
    radius = rng.uniform(0.05, 0.15, n)
    volume = rng.uniform(0.5, 1.5, n)
    birth = np.linalg.norm(alpha, axis=1) * rng.uniform(0.3, 0.7, n)
    active = rng.uniform(0.0, 10.0, n)
    return pos, alpha, radius, volume, birth, active


def _arrays(pos, alpha, radius, volume, birth, active, maxn):
    out = []
    for a, cols in ((pos, 3), (alpha, 3), (radius, 0), (volume, 0), (birth, 0), (active, 0)):
        b = np.zeros((maxn, 3) if cols else maxn)
        b[:len(a)] = a
        out.append(b)
    return out


@pytest.mark.parametrize("n,seed,stepnum", [(2000, 1, 70.0), (12000, 2, 80.0), (30000, 3, 60.0)])
def test_merge_matches_the_sequential_oracle(n, seed, stepnum):
    pos, alpha, radius, volume, birth, active = _clustered(n, seed)
    active[int(0.8 * n)] = stepnum  # age 0: the outer loop stops there
    w = O.WakeParticles(pos, alpha, radius, volume, birth, active)
    k_ref = O.wake_merge(w, stepnum)
    g = [np.ascontiguousarray(a) for a in (pos, alpha, radius, volume, active)]
    k = capi.wake_merge(g[0], g[1], g[2], g[3], g[4], n, stepnum)
    _say(f"merge n={n} seed={seed}: {k} merges on the GPU, {k_ref} in the oracle")
    assert k == k_ref and k > n // 50
    np.testing.assert_array_equal(g[0], w.pos)
    np.testing.assert_array_equal(g[1], w.alpha)
    np.testing.assert_array_equal(g[3], w.volume)


@pytest.mark.parametrize("n,seed", [(3000, 4), (40000, 5)])
def test_split_matches_the_sequential_oracle(n, seed):
    pos, alpha, radius, volume, birth, active = _clustered(n, seed)
    birth[::3] *= 0.4  # a third have doubled their strength
    stepnum = 60.0
    active[int(0.9 * n)] = stepnum
    maxn = 2 * n
    w = O.WakeParticles(pos, alpha, radius, volume, birth, active, maxn)
    k_ref = O.wake_split(w, stepnum)
    g = _arrays(pos, alpha, radius, volume, birth, active, maxn)
    n1, k = capi.wake_split(*g, n=n, stepnum=stepnum, maxn=maxn)
    _say(f"split n={n}: {k} on the GPU, {k_ref} in the oracle")
    assert k == k_ref > 0 and n1 == w.n == n + k
    for a, b in zip(g, (w.pos, w.alpha, w.radius, w.volume, w.birth, w.active)):
        np.testing.assert_array_equal(a, b)


def test_split_refusals_change_nothing():
    pos, alpha, radius, volume, birth, active = _clustered(500, 6)
    birth[:] = 0.1 * np.linalg.norm(alpha, axis=1)
    g = _arrays(pos, alpha, radius, volume, birth, active, 510)  # room for 10
    before = [a.copy() for a in g]
    with pytest.raises(capi.PawanError) as e:
        capi.wake_split(*g, n=500, stepnum=100.0, maxn=510)
    assert e.value.code == capi.PW_ERR_UNSUPPORTED
    for a, b in zip(g, before):
        np.testing.assert_array_equal(a, b)


def test_relax_the_helical_wake():
    """1M particles; one in 500 has its strength turned by 90 degrees, so it
    merges with its helix neighbours.  The total strength is kept to
    rounding, and the relaxed wake's right-hand side is finite."""
    s = W.helix_wake(1 << 20)
    n = s.n
    pos, alpha = s.pos.copy(), s.alpha.copy()
    rng = np.random.default_rng(9)
    turn = rng.choice(n, size=n // 500, replace=False)
    t = alpha[turn]
    alpha[turn] = np.stack([-t[:, 1], t[:, 0], t[:, 2]], axis=1)
    radius = s.radius.copy()
    volume = s.volume.copy()
    active = np.zeros(n)
    s0 = [math.fsum(alpha[:, c]) for c in range(3)]  # exact sums
    scale = np.abs(alpha).sum()
    t0 = time.perf_counter()
    k = capi.wake_merge(pos, alpha, radius, volume, active, n, 100.0)
    dt = time.perf_counter() - t0
    _say(f"helix N={n}: {k} merges, {dt:.2f} s (with transfers)")
    assert k >= len(turn)
    # only the roundings of the merges' a_i + a_j remain
    assert max(abs(math.fsum(alpha[:, c]) - s0[c]) for c in range(3)) < 1e-14 * scale
    with capi.Solver() as g:
        g.add_wake(pos, alpha, radius, volume)
        u, w = g.solve("fp32c")
    assert np.all(np.isfinite(u)) and np.all(np.isfinite(w))


@pytest.mark.parametrize("split", [False, True])
def test_relax_through_the_class_api(tmp_path, split):
    """__interaction::relax over a wake copied into MAXNUMPARTICLES rows, after
    and before a right-hand side on the device: the wake arrays are the
    oracle's split-then-merge bit for bit, and the right-hand side of the
    relaxed wake is the oracle's to 1e-12 (with splits the device copy is
    rebuilt; without, it is updated in place)."""
    n, stepnum, nu = 3000, 60.0, 1e-4
    pos, alpha, radius, volume, birth, active = _clustered(n, 11)
    if split:
        birth[::3] *= 0.4
    else:
        birth[:] = np.linalg.norm(alpha, axis=1)  # nothing has doubled
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        f.write(np.int64(n).tobytes() + np.float64(stepnum).tobytes())
        for a in (pos, alpha, radius, volume, birth, active):
            f.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())
    r = subprocess.run([DRIVER, str(inp), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    w = O.WakeParticles(pos, alpha, radius, volume, birth, active, maxn=100000)
    ns = O.wake_split(w, stepnum)
    nm = O.wake_merge(w, stepnum)
    buf = np.fromfile(out, dtype=np.float64)
    m = int(buf[:1].view(np.int64)[0])
    o = 1
    got = {}
    for name, cols in (("pos", 3), ("alpha", 3), ("volume", 1), ("birth", 1), ("u", 3), ("w", 3)):
        got[name] = buf[o:o + cols * m].reshape(m, cols) if cols > 1 else buf[o:o + m]
        o += cols * m
    _say(f"class API relax: {ns} splits, {nm} merges, N {n} -> {m}")
    assert m == w.n and (ns > 0) == split and nm > 0
    np.testing.assert_array_equal(got["pos"], w.pos[:m])
    np.testing.assert_array_equal(got["alpha"], w.alpha[:m])
    np.testing.assert_array_equal(got["volume"], w.volume[:m])
    np.testing.assert_array_equal(got["birth"], w.birth[:m])
    s = W.ParticleSystem(off=np.array([0, m], dtype=np.int64), pos=w.pos[:m].copy(), alpha=w.alpha[:m].copy(),
                         radius=w.radius[:m].copy(), volume=w.volume[:m].copy(), nu=nu, coupled=True, name="relaxed")
    ru, rw = O.solve(s)
    from conftest import normwise
    eu, ew = normwise(got["u"], ru), normwise(got["w"], rw)
    _say(f"  rhs of the relaxed wake vs oracle: e_u={eu:.2e} e_w={ew:.2e}")
    assert eu < 1e-12 and ew < 1e-12
