"""pw_velocity_at_noalloc, the point velocity that allocates nothing (its
points and velocities pass through the library's static staging array),
against pw_velocity_at on the same context and the oracle: two calls in a
row, the second with fewer points, then the limits."""
import numpy as np
import pytest

from conftest import normwise
from pawan_amd import capi, workloads as W

pytestmark = pytest.mark.gpu

MAX_POINTS = 65536


def _say(msg):
    print(msg, flush=True)


def test_noalloc_point_velocity_matches_the_allocating_entry(oracle):
    s = W.leapfrog_64k(nst=256)  # 4,096 particles, two coupled wakes
    rng = np.random.default_rng(21)
    big = np.concatenate([rng.uniform(-1.6, 1.6, size=(MAX_POINTS - 96, 3)) * [1, 1, 0.6],
                          s.pos[:96] + 1e-3])
    small = rng.uniform(-1.2, 1.2, size=(777, 3))
    with capi.Solver(s) as g:
        # the particles reach the device on the context's first device call
        with pytest.raises(capi.PawanError) as e:
            g.velocity_at_noalloc(small)
        assert e.value.code == capi.PW_ERR_STATE
        g.solve("fp64")
        for pts in (big, small):
            _say(f"leapfrog N={s.n}: {len(pts)} points through the static staging array")
            u0 = g.velocity_at_noalloc(pts)
            u1 = g.velocity_at(pts)
            np.testing.assert_array_equal(u0, u1)  # the same kernel on the same data
            sample = np.linspace(0, len(pts) - 1, min(len(pts), 400)).astype(np.int64)
            ref = oracle.velocity_at(s, pts[sample])
            e_u = normwise(u0[sample], ref)
            _say(f"  vs oracle: {e_u:.2e}")
            assert e_u < 1e-12
        # a second context on the same GPU shares the array
        with capi.Solver(W.ring_1k()) as h:
            h.solve("fp64")
            ur = h.velocity_at_noalloc(small)
            np.testing.assert_array_equal(ur, h.velocity_at(small))
        # beyond the array: an argument error, nothing launched
        with pytest.raises(capi.PawanError) as e:
            g.velocity_at_noalloc(np.zeros((MAX_POINTS + 1, 3)))
        assert e.value.code == capi.PW_ERR_ARG
        # and the context still answers after the refused call
        np.testing.assert_array_equal(g.velocity_at_noalloc(small), g.velocity_at(small))
