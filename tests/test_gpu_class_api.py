"""The C++ class API (__parallel / __rk4 / __rk3 / integrate) driven by
bin/pawan_run, a driver written like the reference's validation main
(src/pawan.cpp:140-177).  Its .wake and .diagnosis files are parsed with the
layout of the reference's readers (scripts/readWake.py:89-133,
scripts/readDiagnostics.py:69-83) and compared with the oracle."""
import os
import struct
import subprocess

import numpy as np
import pytest

from conftest import normwise
from pawan_amd import workloads as W

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "pawan_amd", "bin", "pawan_run")


def write_input(path, s):
    with open(path, "wb") as f:
        f.write(struct.pack("q", s.nw))
        for k in range(s.nw):
            a, b = int(s.off[k]), int(s.off[k + 1])
            f.write(struct.pack("q", b - a))
            for arr in (s.pos[a:b], s.alpha[a:b], s.radius[a:b], s.volume[a:b]):
                f.write(np.ascontiguousarray(arr, dtype=np.float64).tobytes())


def read_wake(path):
    """Records of [t, nWake, maxN (int + 4 pad bytes), per wake N, pos,
    vorticity, radius, active, volume, birthstrength] (readWake.py:89-133)."""
    buf = open(path, "rb").read()
    recs, o = [], 0
    while o < len(buf):
        t, = struct.unpack_from("d", buf, o); o += 8
        nw, = struct.unpack_from("q", buf, o); o += 8
        maxn, = struct.unpack_from("i", buf, o); o += 8  # int + 4 skipped bytes
        wakes = []
        for _ in range(nw):
            n, = struct.unpack_from("q", buf, o); o += 8
            pos = np.frombuffer(buf, "d", 3 * maxn, o).reshape(maxn, 3); o += 24 * maxn
            vor = np.frombuffer(buf, "d", 3 * maxn, o).reshape(maxn, 3); o += 24 * maxn
            o += 8 * maxn * 4  # radius, active, volume, birthstrength
            wakes.append((n, pos[:n].copy(), vor[:n].copy()))
        recs.append((t, wakes))
    return recs


def read_diagnosis(path):
    """nu, then records [t, Omega(3), I(3), A(3), H, E, EF, KE, KEF, Zc]
    (readDiagnostics.py:69-83)."""
    d = np.fromfile(path, dtype=np.float64)
    return d[0], d[1:].reshape(-1, 16)


@pytest.mark.parametrize("scheme", ["rk4", "rk3"])
def test_driver_matches_oracle(tmp_path, oracle, scheme):
    s = W.leapfrog_64k(nst=32)  # two coupled wakes, 512 particles, nu = 1e-4
    inp = tmp_path / "in.bin"
    write_input(inp, s)
    env = dict(os.environ, PAWAN_DATA_DIR=str(tmp_path))
    nsteps, dt = 6, 0.002
    print(f"pawan_run leapfrog N={s.n} {nsteps} x {scheme}", flush=True)
    r = subprocess.run([EXE, str(inp), "lf", str(nsteps * dt), str(nsteps), scheme, "fp64",
                        repr(s.nu), "1"], env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    recs = read_wake(tmp_path / "lfWake.wake")
    assert len(recs) == nsteps + 1
    t, wakes = recs[-1]
    assert t == pytest.approx(nsteps * dt, rel=1e-15)
    pos = np.concatenate([p for _, p, _ in wakes])
    vor = np.concatenate([v for _, _, v in wakes])
    ref = oracle.step(s, scheme, dt, nsteps)
    assert normwise(pos, ref.pos) < 1e-12
    assert normwise(vor, ref.alpha) < 1e-12
    nu, diag = read_diagnosis(tmp_path / "lfWake.diagnosis")
    assert nu == s.nu and diag.shape == (nsteps + 1, 16)
    d_ref = oracle.diagnose(ref)
    np.testing.assert_allclose(diag[-1, 13:15], d_ref[12:14], rtol=1e-12)   # KE, KEF
    np.testing.assert_allclose(diag[-1, 4:7], d_ref[3:6], rtol=1e-12, atol=1e-15)  # linear impulse
    assert diag[-1, 15] == pytest.approx(d_ref[14], rel=1e-12)  # centroid of the last wake


def test_driver_runs_the_tree_path(tmp_path):
    """The same driver with PW_LEARNED: it trains the network on its own
    initial state through the C-ABI, integrates, and stays close to the
    FP64 run of the same driver."""
    s = W.helix_wake(16384, seed=5)
    inp = tmp_path / "in.bin"
    write_input(inp, s)
    env = dict(os.environ, PAWAN_DATA_DIR=str(tmp_path), PAWAN_FF_SAMPLES="60000", PAWAN_FF_STEPS="2000")
    out = {}
    for prec in ("fp64", "learned"):
        r = subprocess.run([EXE, str(inp), prec, "0.002", "2", "rk4", prec, repr(s.nu), "0"], env=env,
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        print(r.stderr.strip(), flush=True)
        t, wakes = read_wake(tmp_path / f"{prec}Wake.wake")[-1]
        out[prec] = (wakes[0][1], wakes[0][2])
    disp = out["fp64"][0] - s.pos
    assert normwise(out["learned"][0] - s.pos, disp) < 1e-2
    assert normwise(out["learned"][1], out["fp64"][1]) < 1e-3


def test_driver_resolves_a_vring(tmp_path, oracle):
    """__vring -> __interaction -> __resolve::rebuild -> copy into a __wake ->
    __rk4, as the reference's main builds its rings (src/pawan.cpp:57-84)."""
    R, core, nl, nazi, sig = 1.0, 0.1, 2, 30, 0.15
    env = dict(os.environ, PAWAN_DATA_DIR=str(tmp_path))
    nsteps, dt, nu = 3, 0.002, 1e-4
    print(f"pawan_run vring L={nl} nAzi={nazi}, {nsteps} x rk4", flush=True)
    r = subprocess.run([EXE, f"vring:{R},{core},{nl},{nazi},{sig}", "vr", str(nsteps * dt), str(nsteps),
                        "rk4", "fp64", repr(nu), "0"], env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    recs = read_wake(tmp_path / "vrWake.wake")
    _, (pos, alpha, radius, volume) = oracle.vring_resolve(R, core, nl, nazi, sig)[1:]
    n0, p0, v0 = recs[0][1][0]
    assert n0 == len(pos)
    # the two generators call libm's sin/cos in their own ways (a compiler
    # may merge them into sincos): positions agree to an ulp
    np.testing.assert_allclose(p0, pos, rtol=0, atol=4e-16)
    assert normwise(v0, alpha) < 1e-12
    s = W.ParticleSystem(off=np.array([0, len(pos)], dtype=np.int64), pos=pos, alpha=alpha, radius=radius,
                         volume=volume, nu=nu)
    ref = oracle.step(s, "rk4", dt, nsteps)
    _, p1, v1 = recs[-1][1][0]
    print(f"  e_pos={normwise(p1, ref.pos):.2e} e_alpha={normwise(v1, ref.alpha):.2e}", flush=True)
    assert normwise(p1, ref.pos) < 1e-12
    assert normwise(v1, ref.alpha) < 1e-11
    # the .wakeinfluence file of the resolve (resolve.cpp:16-35): the initial
    # state, then one record per unit cell strength, t = 0, 1, ..., m
    m = (2 * nl + 1) ** 2
    infl = read_wake(tmp_path / "vr_vringWake.wakeinfluence")
    assert [t for t, _ in infl] == [float(i) for i in range(m + 1)]
    psi = np.repeat(np.arange(nazi) * 2 * np.pi / nazi, m)
    cell = np.tile(np.arange(m), nazi)
    for i in (1, 2, m):
        n_i, p_i, v_i = infl[i][1][0]
        assert n_i == len(pos)
        np.testing.assert_allclose(p_i, pos, rtol=0, atol=4e-16)
        on = (cell == i - 1).astype(float)
        np.testing.assert_allclose(v_i[:, 0], -np.sin(psi) * on, rtol=0, atol=1e-15)
        np.testing.assert_allclose(v_i[:, 1], np.cos(psi) * on, rtol=0, atol=1e-15)
        assert np.all(v_i[cell != i - 1] == 0.0)
    assert np.all(infl[0][1][0][2] == 0.0)  # the ring starts without vorticity
