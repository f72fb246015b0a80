"""Profiling driver: right-hand sides of one workload, plus a plain copy
kernel of known size, for rocprofv3 counter runs.

    rocprofv3 --pmc FETCH_SIZE --kernel-trace -d OUT -- \
        python -m pawan_amd.profile_rhs --workload helix --n 1048576 --precision fp32c

The copy (torch's elementwise copy of --copy-mb MiB, float32) reads and
writes a known number of bytes, so the counters' KiB per dispatch can be
scaled to bytes: factor = counter(copy) / (copy bytes / 1024).
"""
from __future__ import annotations

import argparse
import sys
import time


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="helix", choices=["helix", "batched_256", "leapfrog_64k", "ring_1k"])
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--precision", default="fp32c", choices=["fp64", "fp64_ref", "fp32c", "learned"])
    ap.add_argument("--reps", type=int, default=1)
    ap.add_argument("--copy-mb", type=int, default=256)
    ap.add_argument("--theta", type=float, default=0.5)
    ap.add_argument("--tol", type=float, default=0.0,
                    help="tree path: absolute far-field error tolerance (0: opening angle only)")
    ap.add_argument("--config", default="",
                    help="pair-kernel launch shape 'parts,persistent,waves8' (e.g. 8,1,1); default: untuned")
    ap.add_argument("--kernel", default="gaussian", choices=["gaussian", "highorder"],
                    help="smoothing kernel (highorder: FP64 precisions only)")
    args = ap.parse_args()

    import torch

    from pawan_amd import capi, workloads

    nbytes = args.copy_mb << 20
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    print(f"[profile_rhs] copy of {nbytes} bytes done", flush=True)
    if args.workload == "helix":
        system = workloads.helix_wake(args.n)
    else:
        system = workloads.WORKLOADS[args.workload]["make"]()
    with capi.Solver(system) as g:
        if args.kernel == "highorder":
            g.set_kernel(capi.PW_KERNEL_HIGHORDER)
        if args.precision == "learned":
            g.ff_set_params(args.theta, args.tol, use_mlp=False)
        elif args.config:
            parts, pers, w8 = (int(v) for v in args.config.split(","))
            g.set_pair_config(args.precision, parts, bool(pers), bool(w8))
        for k in range(args.reps):
            t = time.perf_counter()
            g.solve(args.precision)
            print(f"[profile_rhs] {system.name} N={system.n} {args.precision} rhs {k + 1}/{args.reps}: "
                  f"{time.perf_counter() - t:.3f} s", flush=True)
        if args.precision == "learned":
            print(f"[profile_rhs] tree: {g.ff_info()}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
