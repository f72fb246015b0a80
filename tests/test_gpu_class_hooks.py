"""The reference's overridable hooks of __interaction on the device: the
per-wake-block entry points (pw_interact_wakes, pw_vorticity_field,
pw_diagnose_wakes) against the oracle, and a C++ subclass that replaces
interact(__wake*) and calculateKineticEnergy(__wake*) as the reference's
__parallel replaces __interaction's (parallel.h:48-70), run through
solve(), diagnose() and __rk4::integrate."""
import os
import struct
import subprocess

import numpy as np
import pytest

from conftest import normwise
from pawan_amd import capi, workloads as W

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOOKS = os.path.join(ROOT, "tests", "cpp", "bin", "hooks_driver")


def _say(msg):
    print(msg, flush=True)


def _two_wakes():
    return W.leapfrog_64k(nst=32)  # 512 particles in two coupled wakes, nu = 1e-4


def test_interact_blocks_sum_to_the_system(oracle):
    s = _two_wakes()
    a, b = int(s.off[1]), s.n
    u0, w0 = oracle.solve(s.wake(0))
    u1, w1 = oracle.solve(s.wake(1))
    uu, ww = oracle.solve(s)
    with capi.Solver(s) as g:
        us, ws = g.interact_wakes("fp64", [(0, 0)])
        uc, wc = g.interact_wakes("fp64", [(0, 1)])
        ua, wa = g.interact_wakes("fp64", [(0, 0), (1, 1), (0, 1)])
        u32, _ = g.interact_wakes("fp32c", [(1, 1), (0, 1)])
        with pytest.raises(capi.PawanError) as e:
            g.interact_wakes("fp64", [(1, 0)])
        assert e.value.code == capi.PW_ERR_ARG
    _say(f"self block e_u={normwise(us[:a], u0):.2e}; all blocks e_u={normwise(ua, uu):.2e}")
    assert normwise(us[:a], u0) < 1e-12 and normwise(ws[:a], w0) < 1e-12
    assert np.all(us[a:] == 0) and np.all(ws[a:] == 0)
    cross_u = uu - np.concatenate([u0, u1])
    cross_w = ww - np.concatenate([w0, w1])
    assert normwise(uc, cross_u) < 1e-12 * np.max(np.linalg.norm(uu, axis=1)) / np.max(np.linalg.norm(cross_u, axis=1))
    assert normwise(ua, uu) < 1e-12 and normwise(wa, ww) < 1e-12
    ref32 = uu - np.concatenate([u0, np.zeros_like(u1)])
    assert normwise(u32, ref32) < 1e-6


def test_vorticity_field_against_numpy():
    """influence(): sum_s ZETASIG(|x_t - x_s|, sigma_ts) a_s plus
    ZETASIG(0, s_t) a_t (interaction_utils.h:626-703), evaluated here in
    numpy, for the whole system and for one cross block."""
    s = _two_wakes()
    x, al, r = s.pos, s.alpha, s.radius
    d = np.linalg.norm(x[:, None, :] - x[None, :, :], axis=-1)
    sig = np.sqrt(0.5 * (r[:, None] ** 2 + r[None, :] ** 2))
    np.fill_diagonal(sig, r)
    Z = np.exp(-0.5 * (d / sig) ** 2) / sig ** 3 / (2 * np.pi) ** 1.5
    full = Z @ al
    a = int(s.off[1])
    cross = np.zeros_like(full)
    cross[:a] = Z[:a, a:] @ al[a:]
    cross[a:] = Z[a:, :a] @ al[:a]
    with capi.Solver(s) as g:
        f_all = g.vorticity_field()
        f_cross = g.vorticity_field([(0, 1)])
        f_sum = g.vorticity_field([(0, 0), (1, 1), (0, 1)])
    _say(f"field e={normwise(f_all, full):.2e} cross e={normwise(f_cross, cross):.2e}")
    assert normwise(f_all, full) < 1e-13 and normwise(f_sum, full) < 1e-13
    assert normwise(f_cross, cross) < 1e-12


def test_diagnose_blocks_sum_to_the_system(oracle):
    s = _two_wakes()
    with capi.Solver(s) as g:
        tot = g.diagnose()[:15]
        d00, d11, d01 = g.diagnose_wakes(0, 0), g.diagnose_wakes(1, 1), g.diagnose_wakes(0, 1)
    ref0 = oracle.diagnose(s.wake(0))
    np.testing.assert_allclose(d00[12:14], ref0[12:14], rtol=1e-12)       # KE, KEF of wake 0 alone
    np.testing.assert_allclose(d00[3:6], ref0[3:6], rtol=1e-12, atol=1e-15)
    summed = d00 + d11 + d01
    np.testing.assert_allclose(summed[9:14], tot[9:14], rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(summed[0:9], tot[0:9], rtol=1e-12, atol=1e-15)
    assert np.all(d01[0:9] == 0) and np.all(d01[14:] == 0)
    assert d11[14] / d11[15] == pytest.approx(tot[14], rel=1e-12)  # Zc of the last wake


def _write_input(path, s):
    with open(path, "wb") as f:
        f.write(struct.pack("q", s.nw))
        for k in range(s.nw):
            a, b = int(s.off[k]), int(s.off[k + 1])
            f.write(struct.pack("q", b - a))
            for arr in (s.pos[a:b], s.alpha[a:b], s.radius[a:b], s.volume[a:b]):
                f.write(np.ascontiguousarray(arr, dtype=np.float64).tobytes())


def test_subclass_overriding_the_hooks(tmp_path, oracle):
    s = _two_wakes()
    inp = tmp_path / "in.bin"
    _write_input(inp, s)
    dt = 0.002
    env = dict(os.environ, PAWAN_DATA_DIR=str(tmp_path))
    name = "hk"  # the driver's files: hk.rates, hk.diag (cwd) and hkWake.wake (__io)
    r = subprocess.run([HOOKS, str(inp), name, repr(dt), repr(s.nu)], env=env, capture_output=True, text=True,
                       timeout=120, cwd=str(tmp_path))
    assert r.returncode == 0, r.stderr
    # integrate() prints its progress; the driver's own last line counts the hook calls
    last = [l for l in r.stdout.splitlines() if l.startswith("self_calls")][-1]
    _say(last)
    calls = dict(zip(last.split()[0::2], map(int, last.split()[1::2])))
    # solve() once, then four stages of the step on the host path, two
    # wakes each; diagnose() once per wake
    assert calls == {"self_calls": 10, "ke_calls": 2}
    rates = np.fromfile(str(tmp_path / "hk.rates"), dtype=np.float64)
    u, w = [], []
    o = 0
    for k in range(s.nw):
        n = int(s.off[k + 1] - s.off[k])
        u.append(rates[o:o + 3 * n].reshape(n, 3)); o += 3 * n
        w.append(rates[o:o + 3 * n].reshape(n, 3)); o += 3 * n
    ru, rw = oracle.solve(s)
    eu, ew = normwise(np.concatenate(u), ru), normwise(np.concatenate(w), rw)
    _say(f"host self pairs + device cross pairs vs oracle: e_u={eu:.2e} e_w={ew:.2e}")
    assert eu < 1e-12 and ew < 1e-12
    d = np.fromfile(str(tmp_path / "hk.diag"), dtype=np.float64)
    ref = oracle.diagnose(s)
    ke_self = sum(oracle.diagnose(s.wake(k))[12] for k in range(s.nw))
    assert d[12] == pytest.approx(ref[12] + ke_self, rel=1e-12)   # the override doubled the self parts
    assert d[13] == pytest.approx(ref[13], rel=1e-12)
    np.testing.assert_allclose(d[3:6], ref[3:6], rtol=1e-12, atol=1e-15)
    # one RK4 step through the reference's host sequence (the overridden
    # block cannot run inside a device step)
    buf = open(os.path.join(str(tmp_path), "hkWake.wake"), "rb").read()
    # last record: t, nWake, maxN (8 bytes), then per wake N and N x 3 positions
    rec = len(buf) // 2
    o = rec + 8 + 8 + 8
    pos = []
    for k in range(s.nw):
        n, = struct.unpack_from("q", buf, o); o += 8
        pos.append(np.frombuffer(buf, "d", 3 * n, o).reshape(n, 3)); o += 8 * n * 10
    st = oracle.step(s, "rk4", dt, 1)
    ep = normwise(np.concatenate(pos) - s.pos, st.pos - s.pos)
    _say(f"one RK4 step vs oracle: e_dx={ep:.2e}")
    assert ep < 1e-11
