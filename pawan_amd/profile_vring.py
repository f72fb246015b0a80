"""Profiling driver: strength resolution of vortex rings (__resolve::rebuild
on a __vring), three resolves each, for rocprofv3 kernel traces.

    rocprofv3 --kernel-trace --stats -d OUT -- python -m pawan_amd.profile_vring

Rings: the reference's own (R = 1, core 0.1, 3 layers, 49 stations,
sigma 0.1924; 49 cells, 2,401 particles) and the largest the entry point
takes (15 layers, 961 cells, 256 stations: 246,016 particles).
"""
from __future__ import annotations

import sys
import time

import numpy as np

RINGS = ((3, 49, 0.1924), (15, 256, 0.02))


def main() -> int:
    from pawan_amd import capi

    for nl, nazi, sig in RINGS:
        pos, rad, vol = capi.vring_particles(1.0, 0.1, nl, nazi, sig)
        with capi.Solver() as g:
            g.add_wake(pos, np.zeros_like(pos), rad, vol)
            for k in range(3):
                t = time.perf_counter()
                g.vring_resolve(0, nl, nazi, 1.0, 0.1)
                print(f"[profile_vring] vring L={nl} nAzi={nazi} N={len(pos)} resolve {k + 1}/3: "
                      f"{time.perf_counter() - t:.4f} s", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
