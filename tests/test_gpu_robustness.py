"""Error paths and switches of the direct sums on the GPU: coincident
particles (the reference's kernel gives 0/0 there, gaussian.h:53-59),
viscosity changed between graph-replayed steps, and what the compensated
FP32 accumulation buys over a plain FP32 sum."""
import os

import numpy as np
import pytest

from conftest import normwise
from pawan_amd import capi, workloads as W

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _say(msg):
    print(msg, flush=True)


def _coincident_ring():
    s = W.ring_1k()
    s.pos[7] = s.pos[3]  # two distinct particles at one point
    return s


@pytest.mark.parametrize("prec", ["fp64", "fp64_ref"])
def test_coincident_particles_are_reported(prec):
    """The FP64 paths evaluate the reference's kernel, which is 0/0 at
    rho = 0: the rates come out non-finite and the call returns
    PW_ERR_NONFINITE (solve and step); the context stays usable."""
    s = _coincident_ring()
    _say(f"coincident particles 3 and 7 of ring_1k, {prec}")
    with capi.Solver(s) as g:
        with pytest.raises(capi.PawanError) as e:
            g.solve(prec)
        assert e.value.code == capi.PW_ERR_NONFINITE
        with pytest.raises(capi.PawanError) as e:
            g.step("rk4", 0.005, prec, 1)
        assert e.value.code == capi.PW_ERR_NONFINITE
        ok = W.ring_1k()
        g.set_wake_state(0, ok.pos, ok.alpha)
        u, w = g.solve(prec)
        assert np.all(np.isfinite(u)) and np.all(np.isfinite(w))


def test_coincident_particles_fp32c_take_the_limit(oracle):
    """The FP32 near-field form (positive series in rho_bar^2) is exact at
    rho = 0: the pair adds no velocity and q(0) (a_t x a_s) plus the PSE
    exchange to the stretching.  It matches the __float128 oracle on the
    same system with one of the two particles moved by 1e-9 sigma (where the
    FP64 formula's cancellation, ~eps / rho_bar^2, would cost it 1% of that
    pair's stretching)."""
    s = _coincident_ring()
    ref = s.copy()
    ref.pos[7] = ref.pos[3] + np.array([1e-10, 0.0, 0.0])
    ru, rw = oracle.solve_q128(ref)
    with capi.Solver(s) as g:
        u, w = g.solve("fp32c")
    eu, ew = normwise(u, ru), normwise(w, rw)
    _say(f"fp32c at coincidence vs __float128 oracle displaced by 1e-9 sigma: e_u={eu:.2e} e_w={ew:.2e}")
    assert eu < 1e-6 and ew < 1e-5


def test_viscosity_change_between_graph_replays(oracle):
    """nu is a kernel argument of the captured step: changing it drops the
    graphs, so the next replayed step uses the new value (bit-identical to
    eager steps, and to the oracle at 1e-12)."""
    s = W.leapfrog_64k(nst=64)  # 1,024 particles, two coupled wakes
    dt, nu0, nu1 = 2e-3, 1e-4, 5e-3
    s.nu = nu0
    with capi.Solver(s) as gr, capi.Solver(s) as ea:
        gr.set_graph(True)
        for g in (gr, ea):
            g.set_nu(nu0)
            g.step("rk4", dt, "fp64", 2)
            g.set_nu(nu1)
            g.step("rk4", dt, "fp64", 2)
        pg, ag = gr.state()
        pe, ae = ea.state()
    np.testing.assert_array_equal(pg, pe)
    np.testing.assert_array_equal(ag, ae)
    mid = oracle.step(s, "rk4", dt, 2)
    mid.nu = nu1
    end = oracle.step(mid, "rk4", dt, 2)
    _say(f"graph after set_nu vs oracle: pos {normwise(pg - s.pos, end.pos - s.pos):.2e} "
         f"alpha {normwise(ag, end.alpha):.2e}")
    assert normwise(pg - s.pos, end.pos - s.pos) < 1e-11
    assert normwise(ag, end.alpha) < 1e-12


def test_compensation_holds_where_a_plain_fp32_sum_fails():
    """1M-particle helical wake, the fixture's targets (FP64 oracle values):
    summed in one plain FP32 accumulator per target, a million pair terms
    lose the 1e-6 velocity bar by two orders of magnitude; the chunked
    two-sum accumulation of PW_FP32C holds it."""
    g = dict(np.load(os.path.join(HERE, "helix_1m.npz"), allow_pickle=False))
    s = W.helix_wake(1 << 20)
    idx = g["targets"]
    with capi.Solver(s) as gpu:
        u_c, w_c = gpu.solve("fp32c")
        gpu.set_compensation(False)
        u_p, w_p = gpu.solve("fp32c")
        gpu.set_compensation(True)
        u_c2, _ = gpu.solve("fp32c")
    ec, ep = normwise(u_c[idx], g["vel"]), normwise(u_p[idx], g["vel"])
    _say(f"helix 1M velocity vs FP64 oracle: compensated {ec:.2e}, plain FP32 {ep:.2e} "
         f"(stretching {normwise(w_c[idx], g['retvor']):.2e} / {normwise(w_p[idx], g['retvor']):.2e})")
    assert ec < 1e-6
    assert ep > 1e-5
    np.testing.assert_array_equal(u_c2, u_c)  # switching back restores the same sums


def test_valu_ceilings_are_ordered():
    """The issue-rate probe runs every form and reports a positive rate: FP32
    FMAs on vector registers issue at least as fast as FP64 FMAs, and the CU
    count is the device's."""
    rates = {}
    for form in capi.VALU_FORMS:
        rates[form], cus = capi.valu_ceiling(form, 16)
        assert rates[form] > 0.0
    assert cus >= 1
    assert rates["fp32_fma"] >= 0.9 * rates["fp64_fma"]
    print(f"valu ceilings (wave instructions per CU per ns, 16 waves/CU): {rates}, CUs {cus}")
