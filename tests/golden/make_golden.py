"""Writes tests/golden/*.npz: small input/output vectors of the oracle on the
benchmark workloads (seeded generators of pawan_amd.workloads).

The reference holds no test vectors of its own for this path (it has no
tests), so these fixtures are the oracle's outputs, committed so that the
oracle, the generators and the GPU path are all compared against the same
numbers on every machine.  Regenerate with `python tests/golden/make_golden.py`
(takes about a minute on 8 cores).
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from pawan_amd import workloads as W  # noqa: E402


def inputs(s, rows=None) -> dict:
    """Input rows of the system, so a test can check that the generator
    still produces the same particles (to rounding) on its machine."""
    rows = np.arange(s.n) if rows is None else np.asarray(rows)
    return dict(in_rows=rows, in_pos=s.pos[rows], in_alpha=s.alpha[rows], in_radius=s.radius[rows],
                in_volume=s.volume[rows], in_off=s.off, in_nu=np.float64(s.nu))


def sample(n: int, count: int) -> np.ndarray:
    return np.unique(np.linspace(0, n - 1, count).astype(np.int64))


def main() -> None:
    out = {}
    # single ring, N=1000: full rates, two RK4 steps, diagnostics
    r = W.ring_1k()
    u, w = O.solve(r)
    st = O.step(r, "rk4", 0.005, 2)
    out["ring_1k"] = dict(**inputs(r), vel=u, retvor=w, rk4x2_pos=st.pos, rk4x2_alpha=st.alpha,
                          diag=O.diagnose(r), diag_rk4x2=O.diagnose(st))
    # leapfrog, N=65,536 (two coupled wakes): sampled targets
    lf = W.leapfrog_64k()
    t = sample(lf.n, 64)
    u, w = O.rates_targets(lf, t)
    out["leapfrog_64k"] = dict(**inputs(lf, sample(lf.n, 256)), targets=t, vel=u, retvor=w)
    # small leapfrog, 3 RK3 steps with viscosity
    sl = W.leapfrog_64k(nst=16)
    st = O.step(sl, "rk3", 0.01, 3)
    out["leapfrog_256_rk3"] = dict(**inputs(sl), pos=st.pos, alpha=st.alpha)
    # batched rings (6 independent wakes): per-wake rates
    b = W.batched_rings(nwakes=6, nmin=64, nmax=700, seed=11)
    vel, ret = [], []
    for k in range(b.nw):
        u, w = O.solve(b.wake(k))
        vel.append(u)
        ret.append(w)
    out["batched_6"] = dict(**inputs(b), vel=np.concatenate(vel), retvor=np.concatenate(ret))
    # helical wake, N=1,048,576: sampled targets against all sources
    hx = W.helix_wake(1 << 20)
    t = sample(hx.n, 24)
    u, w = O.rates_targets(hx, t)
    out["helix_1m"] = dict(**inputs(hx, sample(hx.n, 512)), targets=t, vel=u, retvor=w)
    for name, d in out.items():
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **d)
        print("wrote", name)


if __name__ == "__main__":
    main()
