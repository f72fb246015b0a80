"""Vortex-ring strength resolution (__resolve::rebuild, resolve.cpp:12-63)
on the GPU against the oracle: the ring's particles, the influence matrix
built on the device, the pivoted LU solve and the strengths written into
the wake."""
import numpy as np
import pytest

from conftest import normwise
from pawan_amd import capi, workloads as W

pytestmark = pytest.mark.gpu

R, CORE, L, NAZI, SIG = 1.0, 0.1, 3, 49, 0.1924   # the reference's own ring (src/pawan.cpp:57)


def _say(msg):
    print(msg, flush=True)


def _check_resolve(oracle, g, wake, s, M, nl=L, nazi=NAZI, sig=SIG):
    s_ref, M_ref, (_, alpha_ref, _, _) = oracle.vring_resolve(R, CORE, nl, nazi, sig)
    ideal = oracle.vring_ideal(R, CORE, nl)
    eM = np.abs(M - M_ref).max() / np.abs(M_ref).max()
    back = np.abs(M_ref @ s - ideal).max() / np.abs(ideal).max()
    fwd = np.abs(s - s_ref).max() / np.abs(s_ref).max()
    _say(f"  M {eM:.2e}  backward {back:.2e}  forward {fwd:.2e}")
    assert eM < 1e-13
    assert back < 1e-9          # cond(M) ~ 1.5e8: the forward error is held to cond x eps
    assert fwd < 1e-6
    _, alpha = g.wake_state(wake)
    m = len(s)
    psi = np.repeat(np.arange(nazi) * 2 * np.pi / nazi, m)
    st = np.tile(s, nazi)
    np.testing.assert_allclose(alpha[:, 0], -np.sin(psi) * st, rtol=1e-14, atol=1e-14 * np.abs(st).max())
    np.testing.assert_allclose(alpha[:, 1], np.cos(psi) * st, rtol=1e-14, atol=1e-14 * np.abs(st).max())
    assert np.all(alpha[:, 2] == 0.0)
    assert normwise(alpha, alpha_ref) < 1e-6


def test_resolve_the_reference_ring(oracle):
    pos, radius, volume = capi.vring_particles(R, CORE, L, NAZI, SIG)
    _say(f"vring L={L} nAzi={NAZI}: {len(pos)} particles")
    with capi.Solver() as g:
        g.add_wake(pos, np.zeros_like(pos), radius, volume)
        s, M = g.vring_resolve(0, L, NAZI, R, CORE)
        _check_resolve(oracle, g, 0, s, M)
        # the resolved ring's right-hand side
        sys_ = W.ParticleSystem(off=np.array([0, len(pos)], dtype=np.int64), pos=pos,
                                alpha=g.wake_state(0)[1].copy(), radius=radius, volume=volume, nu=g.nu)
        u, w = g.solve("fp64")
    ru, rw = oracle.solve(sys_)
    _say(f"  resolved ring rhs: e_u={normwise(u, ru):.2e} e_w={normwise(w, rw):.2e}")
    # the resolved strengths alternate in sign from layer to layer (about
    # +220 at the centre against -49 in the first layer), so the vorticity
    # rate is a difference of terms an order of magnitude larger than itself:
    # held to 1e-11 (measured 3.2e-12), the velocity to 1e-12
    assert normwise(u, ru) < 1e-12 and normwise(w, rw) < 1e-11


def test_ring_behind_another_wake(oracle):
    """The ring as the second wake of a context (offset into the state), with
    the high-order kernel on a smaller ring."""
    nl, nazi, sig = 2, 30, 0.15
    other = W.ring_1k()
    pos, radius, volume = capi.vring_particles(R, CORE, nl, nazi, sig)
    with capi.Solver() as g:
        g.add_wake(other.pos, other.alpha, other.radius, other.volume)
        g.add_wake(pos, np.zeros_like(pos), radius, volume)
        s, M = g.vring_resolve(1, nl, nazi, R, CORE)
        _say("ring as wake 1 (Gaussian)")
        _check_resolve(oracle, g, 1, s, M, nl, nazi, sig)
        np.testing.assert_array_equal(g.wake_state(0)[1], other.alpha)
        g.set_kernel(capi.PW_KERNEL_HIGHORDER)
        s2, M2 = g.vring_resolve(1, nl, nazi, R, CORE)
        _say("ring as wake 1 (high-order)")
        with oracle.smoothing(oracle.HIGHORDER):
            _check_resolve(oracle, g, 1, s2, M2, nl, nazi, sig)


def test_bad_rings_are_errors():
    pos, radius, volume = capi.vring_particles(R, CORE, 1, 8, 0.2)
    with capi.Solver() as g:
        g.add_wake(pos, np.zeros_like(pos), radius, volume)
        for args in ((0, 2, 8), (0, 1, 9), (1, 1, 8), (0, 16, 8)):
            with pytest.raises(capi.PawanError) as e:
                g.vring_resolve(args[0], args[1], args[2], R, CORE)
            assert e.value.code == capi.PW_ERR_ARG
    with pytest.raises(capi.PawanError):
        capi.vring_particles(R, CORE, -1, 8, 0.2)


@pytest.mark.parametrize("nl,nazi,sig", [(8, 16, 0.05), (15, 8, 0.02)])
def test_resolve_rings_beyond_one_block(oracle, nl, nazi, sig):
    """m = 289 and 961 cells: factored column by column over every CU
    (k_lu_pivot, k_lu_update), each lane of the pivot search and of the
    triangular solves handling several rows (condition numbers 5e5 and
    2e3)."""
    pos, radius, volume = capi.vring_particles(R, CORE, nl, nazi, sig)
    _say(f"vring L={nl} nAzi={nazi} ({(2 * nl + 1) ** 2} cells): {len(pos)} particles")
    with capi.Solver() as g:
        g.add_wake(pos, np.zeros_like(pos), radius, volume)
        s, M = g.vring_resolve(0, nl, nazi, R, CORE)
        _check_resolve(oracle, g, 0, s, M, nl, nazi, sig)
