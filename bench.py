#!/usr/bin/env python3
"""RK4 steps/s of the vortex-particle step on one MI355X: FP64 direct sum,
FP32-compensated direct sum, and the learned far field at <= 1e-4 error.

Workload (default): the synthetic rotor-like helical wake, N = 1,048,576
particles (4 contracting tip helices + jittered inboard sheet, sigma = 0.02,
nu = 1.56e-5), RK4 dt = 1e-3.  A "step" is one full RK4 step: four
right-hand sides plus the FP64 stage combination, with the state resident
in HBM.  A direct right-hand side evaluates all N(N-1) ordered pair
interactions (velocity, transposed stretching, PSE diffusion).

`value` is the FP32-compensated direct path (FP32 pair math, compensated
per-target accumulation, FP64 state); the FP64 direct path is timed beside
it, with the FP32C error against it.  Kernel times come from HIP events the
library records on its own stream.

Errors follow one definition throughout: e = max_i |x_i - x_ref,i| /
max_i |x_ref,i| over particles (with the RMS ratio beside it), against the
FP64 direct sum on the same state.

The learned leg trains the far-field network on samples the solver
generates from helical wakes of other seeds, then, for the tree path with
the analytic expansion only and with the network, searches the largest
tolerance whose velocity error against the FP64 direct sum is <= 1e-4 and
times RK4 steps at it.  The drift leg runs 200 RK4 steps of the
leapfrogging rings (N = 65,536) at nu = 1e-4 and nu = 0 with FP64, FP32C and
the tree path side by side: every 10 steps it records each path's
conserved quantities against FP64's at the same time, and the velocity
error of its right-hand side against FP64 on its own state.  The CPU
baseline is the oracle's restatement of the reference's OpenMP loops
(parallel.cpp), timed on a sample of source rows of the same system and
extrapolated to one RK4 step.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
Other measurements (not part of the default run):
  --trajectory          FP32C and FP64 RK4 trajectories of the 1M helical
                        wake side by side (50 steps), compared every 10
  --profile-rhs         right-hand sides of one workload and a plain copy of
                        known size, for rocprofv3 counter runs
  --profile-vring       ring strength resolutions, for rocprofv3 traces
  --probe-valu          vector-issue ceilings of the device per instruction
                        form (the default run records them next to the
                        pair kernels' issue rates)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "RK4 steps/s at N=1M: FP64 direct vs FP32-comp vs learned far-field @1e-4 err"
FLOP_PER_PAIR = 88            # algorithmic FLOP per ordered pair (FMA = 2)
PEAK_FP32_TFLOPS = 157.3      # MI355X published FP32 vector peak
PEAK_FP64_TFLOPS = 78.6       # MI355X published FP64 vector peak
PEAK_BF16_TFLOPS = 2500.0     # MI355X published dense BF16 matrix-core peak
FLOP_PER_MLP_EVAL = 72704     # 16-128-128-128-12 network, FMA = 2


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def time_path(system, precision: str, dt: float, steps: int, warmup: int, tune: bool = True):
    from pawan_amd import capi

    with capi.Solver(system) as g:
        tuned = None
        if tune:
            t = time.perf_counter()
            ms = g.autotune(precision)
            tuned = dict(g.pair_config(precision), rhs_ms=ms)
            log(f"{precision} autotune {time.perf_counter() - t:.1f} s -> {tuned}")
        g.set_graph(True)
        for i in range(warmup):
            t = time.perf_counter()
            g.step("rk4", dt, precision, 1)
            g.synchronize()
            log(f"{precision} warmup step {i + 1}/{warmup}: {time.perf_counter() - t:.3f} s")
        g.synchronize()
        g.set_profiling(True)
        t0 = time.perf_counter()
        for i in range(steps):
            g.step("rk4", dt, precision, 1)
            if steps > 1:
                g.synchronize()
                log(f"{precision} step {i + 1}/{steps}: {time.perf_counter() - t0:.3f} s cumulative")
        g.synchronize()
        wall = time.perf_counter() - t0
        ms, launches = g.pair_timing()
        g.set_profiling(False)
    return wall, ms, launches, tuned


def train_network(n_train: int, samples: int, steps: int, theta: float, seed: int = 11):
    """Samples from seeded training wakes (helical wakes of n_train
    particles and seeds other than the benchmark's, generated here), then
    Adam on the GPU.  Returns (weights, report, seconds)."""
    from pawan_amd import capi, workloads

    t0 = time.perf_counter()
    seeds = [seed, seed + 1, seed + 2, seed + 3]
    with capi.Solver(workloads.helix_wake(n_train, seed=seeds[0])) as tr:
        tr.ff_set_params(theta, 0.0, use_mlp=False)
        for k, sd in enumerate(seeds):
            if k > 0:
                w = workloads.helix_wake(n_train, seed=sd)
                tr.set_wake_state(0, w.pos, w.alpha)
            tr.ff_sample(samples // len(seeds), seed=1000 + sd)
        tgen = time.perf_counter() - t0
        tr.ff_init_weights(seed)
        rep = tr.ff_train(steps, batch=1024, lr=2e-3, seed=seed)
        weights = tr.ff_get_weights()
    rep["sample_seconds"] = tgen
    return weights, rep, time.perf_counter() - t0


def maxnorm(a: np.ndarray, b: np.ndarray) -> float:
    """max_i |a_i - b_i| / max_i |b_i| over particles (rows)."""
    return float(np.max(np.linalg.norm(a - b, axis=1)) / np.max(np.linalg.norm(b, axis=1)))


def rmsnorm(a: np.ndarray, b: np.ndarray) -> float:
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def path_errors(u, w, ref_u, ref_w) -> dict:
    return {"velocity_maxnorm": maxnorm(u, ref_u), "velocity_rms": rmsnorm(u, ref_u),
            "stretching_maxnorm": maxnorm(w, ref_w), "stretching_rms": rmsnorm(w, ref_w)}


def calibrate_tol(g, target_err: float, ref_u: np.ndarray, ref_w: np.ndarray, use_mlp: bool, theta: float,
                  urms: float):
    """Largest tolerance (geometric search) whose velocity error against the
    FP64 direct sum, max_i |du_i| / max_i |u_i|, is at most target_err;
    returns (tol, errors, info) or None."""
    def err_at(tol):
        g.ff_set_params(theta, tol, use_mlp=use_mlp)
        u, w = g.solve("learned")
        return path_errors(u, w, ref_u, ref_w), g.ff_info()

    lo, hi = None, None
    tol = target_err * urms
    best = None
    for _ in range(16):
        e, info = err_at(tol)
        log(f"  {'mlp' if use_mlp else 'md '} tol {tol:.3e}: err {e['velocity_maxnorm']:.3e} "
            f"(rms {e['velocity_rms']:.2e}) near {info['near_pairs']:.3e} far {info['far_evals']:.3e}")
        if e["velocity_maxnorm"] <= target_err:
            lo = tol
            best = (tol, e, info)
            if hi is not None and hi / lo < 1.3:
                break
            tol = tol * 4 if hi is None else np.sqrt(lo * hi)
        else:
            hi = tol
            if lo is not None and hi / lo < 1.3:
                break
            tol = tol / 4 if lo is None else np.sqrt(lo * hi)
    return best


def learned_leg(system, dt: float, steps: int, target_err: float, theta: float, train_n: int,
                train_samples: int, train_steps: int):
    from pawan_amd import capi

    weights, rep, ttrain = train_network(train_n, train_samples, train_steps, theta)
    log(f"network trained in {ttrain:.1f} s: {rep}")
    out = {"training": dict(rep, seconds=ttrain, samples=train_samples, wake_particles=train_n),
           "theta": theta, "target_error": target_err,
           "error_metric": "max_i |du_i| / max_i |u_i| against the FP64 direct sum on the same state"}
    with capi.Solver(system) as g:
        ref_u, ref_w = g.solve("fp64")
        urms = float(np.sqrt((ref_u * ref_u).sum(axis=1).mean()))
        g.ff_set_weights(weights)
        g.ff_set_err_coef(rep["coef_md"], rep["coef_mlp"])
        for use_mlp in (False, True):
            key = "mlp" if use_mlp else "analytic"
            best = calibrate_tol(g, target_err, ref_u, ref_w, use_mlp, theta, urms)
            if best is None:
                out[key] = None
                continue
            tol, err, info = best
            g.ff_set_params(theta, tol, use_mlp=use_mlp)
            g.step("rk4", dt, "learned", 1)  # warm-up
            g.synchronize()
            g.set_profiling(True)
            t0 = time.perf_counter()
            for i in range(steps):
                g.step("rk4", dt, "learned", 1)
            g.synchronize()
            wall = time.perf_counter() - t0
            ms, launches = g.pair_timing()
            stages = g.ff_timing()
            g.set_profiling(False)
            nrhs = max(1, launches)
            rhs_ms = ms / nrhs
            st = {k: v / max(1.0, stages["rhs"]) for k, v in stages.items() if k != "rhs"}
            out[key] = {"far_field": ("monopole + dipole + network residual" if use_mlp else
                                      "monopole + dipole + quadrupole (analytic)"),
                        "steps_per_s": steps / wall, "rhs_ms": rhs_ms,
                        "velocity_error": err["velocity_maxnorm"],
                        "velocity_error_maxnorm": err["velocity_maxnorm"], "errors_vs_fp64": err,
                        "near_pairs_per_s": info["near_pairs"] / (st["near_ms"] * 1e-3),
                        "tol_abs": float(tol), "near_pairs": info["near_pairs"], "far_evals": info["far_evals"],
                        "mlp_evals": info["mlp_evals"], "max_far_nodes": info["max_far"],
                        "max_near_leaves": info["max_near"], "near_core_pairs": info["near_core_pairs"],
                        "stage_ms_per_rhs": st}
            if use_mlp and st["far_mlp_ms"] > 0:
                flop = FLOP_PER_MLP_EVAL * info["mlp_evals"]
                ach = flop / (st["far_mlp_ms"] * 1e-3) / 1e12
                out[key]["roofline_mlp"] = {
                    "bound": "mfma", "kernel": "k_far_mlp", "achieved": ach, "peak": PEAK_BF16_TFLOPS,
                    "unit": "TFLOP/s", "frac": ach / PEAK_BF16_TFLOPS, "traffic": None,
                    "flop_per_launch": flop, "avg_launch_ms": st["far_mlp_ms"],
                    "work": f"{info['mlp_evals']} network evaluations (one per target leaf and far node) "
                            f"x {FLOP_PER_MLP_EVAL} FLOP"}
            log(f"{key}: {out[key]}")
            if use_mlp:
                out["_drift_cfg"] = {"weights": weights, "coef_md": rep["coef_md"], "coef_mlp": rep["coef_mlp"],
                                     "theta": theta, "tol_rel": tol / urms}
            # back to the start state for the next variant
            g.set_wake_state(0, system.pos, system.alpha)
    return out


def _diag_terms(system, u0: np.ndarray) -> dict:
    """Scale of the terms each conserved quantity sums, for normalising
    differences of quantities that are ~0 on these rings."""
    a = np.linalg.norm(system.alpha, axis=1)
    r = np.linalg.norm(system.pos, axis=1)
    return {"omega": float(a.sum()), "angular_impulse": float((r * r * a).sum()),
            "helicity": float((a * np.linalg.norm(u0, axis=1)).sum())}


def _quantities(d: np.ndarray) -> dict:
    return {"omega": d[0:3], "impulse": d[3:6], "angular_impulse": d[6:9], "helicity": d[9:10], "ke": d[12:13]}


def drift_leg(steps: int, every: int, learned_cfg: dict | None, target_err: float, nus=(1e-4, 0.0)):
    """The leapfrogging rings (two coupled wakes, N = 65,536, dt = 0.002),
    advanced `steps` RK4 steps by FP64, FP32C and (when configured) the tree
    path with the network ("learned") and with the expansion to quadrupole
    order ("quadrupole"), side by side, at each viscosity in `nus`.  Every
    `every` steps, for each path:
      * dev_Q: |Q_path(t) - Q_FP64(t)| / S_Q, the distance of its conserved
        quantities from FP64's at the same time (S_Q = |Q(0)| for the linear
        impulse and the kinetic energy, and the size of the summed terms for
        the quantities that are ~0 here: sum|alpha| for the total vorticity,
        sum|x|^2|alpha| for the angular impulse, sum|alpha||u| for the
        helicity);
      * drift_Q: |Q_path(t) - Q_path(0)| / S_Q;
      * the velocity error of its right-hand side against an FP64 right-hand
        side on its own state (max-norm and RMS).
    Each tree path's tolerance is calibrated once on this system to the same
    velocity error target as at 1M."""
    from pawan_amd import capi, workloads

    spec = workloads.WORKLOADS["leapfrog_64k"]
    dt = spec["dt"]
    out = {"workload": "leapfrog_64k", "steps": steps, "dt": dt, "every": every, "runs": []}
    tol_abs = {}
    # path name -> the precision it steps with
    prec = {"fp64": "fp64", "fp32c": "fp32c"}
    if learned_cfg is not None:
        prec.update({"learned": "learned", "quadrupole": "learned"})
    for nu in nus:
        system = spec["make"]()
        system.nu = nu
        run = {"nu": nu, "particles": system.n, "t": [], "paths": {}}
        t0 = time.perf_counter()
        ctx = {p: capi.Solver(system) for p in prec}
        try:
            u0, _ = ctx["fp64"].solve("fp64")
            scale = _diag_terms(system, u0)
            for p, g in ctx.items():
                if prec[p] == "learned":
                    use_mlp = p == "learned"
                    if use_mlp:
                        g.ff_set_weights(learned_cfg["weights"])
                    g.ff_set_err_coef(learned_cfg["coef_md"], learned_cfg["coef_mlp"])
                    if p not in tol_abs:
                        ref_u, ref_w = g.solve("fp64")
                        urms = float(np.sqrt((ref_u * ref_u).sum(axis=1).mean()))
                        best = calibrate_tol(g, target_err, ref_u, ref_w, use_mlp, learned_cfg["theta"], urms)
                        if best is None:
                            raise RuntimeError("no tolerance meets the target on the rings")
                        tol_abs[p] = best[0]
                        out[f"{p}_tol_abs"] = tol_abs[p]
                        out[f"{p}_velocity_error_t0"] = best[1]
                    g.ff_set_params(learned_cfg["theta"], tol_abs[p], use_mlp=use_mlp)
                else:
                    g.set_graph(True)
                run["paths"][p] = {"dev": {}, "drift": {}, "e_u_maxnorm": [], "e_u_rms": []}
            q0 = {p: _quantities(g.diagnose()) for p, g in ctx.items()}
            for k in range(0, steps + 1, every):
                if k > 0:
                    for p, g in ctx.items():
                        g.step("rk4", dt, prec[p], every)
                run["t"].append(k * dt)
                q64 = _quantities(ctx["fp64"].diagnose())
                for p, g in ctx.items():
                    q = _quantities(g.diagnose()) if p != "fp64" else q64
                    rec = run["paths"][p]
                    for name in q:
                        sc = scale.get(name, float(np.linalg.norm(q0["fp64"][name])))
                        rec["dev"].setdefault(name, []).append(float(np.linalg.norm(q[name] - q64[name])) / sc)
                        rec["drift"].setdefault(name, []).append(float(np.linalg.norm(q[name] - q0[p][name])) / sc)
                    if p != "fp64":
                        up, _ = g.solve(prec[p])
                        ur, _ = g.solve("fp64")
                        rec["e_u_maxnorm"].append(maxnorm(up, ur))
                        rec["e_u_rms"].append(rmsnorm(up, ur))
        finally:
            for g in ctx.values():
                g.close()
        for p, rec in run["paths"].items():
            rec["max_dev"] = {k: max(v) for k, v in rec["dev"].items()}
            rec["max_drift"] = {k: max(v) for k, v in rec["drift"].items()}
            if rec["e_u_maxnorm"]:
                rec["max_e_u_maxnorm"] = max(rec["e_u_maxnorm"])
            log(f"drift nu={nu} {p}: max dev {rec['max_dev']} max drift {rec['max_drift']}"
                + (f" max e_u {rec['max_e_u_maxnorm']:.2e}" if rec["e_u_maxnorm"] else ""))
        run["seconds"] = time.perf_counter() - t0
        out["runs"].append(run)
    return out


def pair_kernel_name(precision: str, cfg: dict | None) -> str:
    """The pair kernel a tuned launch runs (the persistent or grid form of it)."""
    if precision == "fp64":
        return "k_pair_fp64s" if (cfg or {}).get("scalar_feed") else "k_pair_fp64"
    cfg = cfg or {}
    if cfg.get("two_per_lane"):
        return "k_pair_fp32s_t2"
    return ("k_pair_fp32s" if cfg.get("scalar_feed") else "k_pair_fp32c") + ("_w8" if cfg.get("waves8") else "")


def valu_ceilings() -> dict:
    """Vector-issue rates this device reaches for the instruction forms the
    pair kernels are made of (pw_valu_ceiling: long chains of one form on
    every CU, the best of 16 and 32 resident waves per CU), in wave
    instructions per CU per nanosecond."""
    from pawan_amd import capi

    out = {"unit": "wave instructions per CU per ns", "probe": "pw_valu_ceiling"}
    for form in capi.VALU_FORMS:
        rates = []
        for w in (16, 32):
            r, cus = capi.valu_ceiling(form, w)
            rates.append(r)
        out[form] = max(rates)
    out["cus"] = cus
    return out


def measured_issue(traffic: dict | None, pairs: int, avg_ms: float, ceilings: dict | None, form: str):
    """The pair kernel's vector-issue rate in the timed launches (its
    instructions per pair from the committed counter runs of the same launch
    shape, times the pairs, over the launch time and the CUs), against the
    measured ceiling of the instruction form it is mostly made of."""
    if not traffic or "issue" not in traffic or not ceilings:
        return None
    vpp = traffic["issue"]["valu_per_pair"]
    rate = vpp * pairs / 64.0 / (avg_ms * 1e6) / ceilings["cus"]
    return {"valu_per_pair": vpp, "per_cu_ns": rate, "ceiling_form": form, "ceiling_per_cu_ns": ceilings[form],
            "frac_of_ceiling": rate / ceilings[form]}


def measured_traffic(workload: str, n: int, precision: str, cfg: dict | None):
    """HBM bytes per launch of the pair kernel from the committed rocprofv3
    counter runs (FETCH_SIZE and WRITE_SIZE in KiB per dispatch, scaled by
    the factor the same counters read for a copy of known size), for the
    workload and launch shape they were measured on; None otherwise."""
    path = os.path.join(ROOT, "profiles", "pair_kernels_helix1m_traffic.txt")
    if workload != "helix" or n != 1 << 20 or not cfg or not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    keys = ("parts", "persistent", "waves8", "scalar_feed", "two_per_lane")
    want = {k: bool(cfg.get(k)) if k != "parts" else int(cfg.get(k, 1)) for k in keys}
    for e in t["entries"]:
        have = {k: (int(e["launch_config"].get(k, 1)) if k == "parts" else bool(e["launch_config"].get(k)))
                for k in keys}
        if e["precision"] == precision and have == want:
            fetch = e["fetch_kib"] * t["fetch_factor"] * 1024.0
            write = e["write_kib"] * t["write_factor"] * 1024.0
            out = {"bytes": fetch + write, "fetch_bytes": fetch, "write_bytes": write, "kernel": e["kernel"],
                   "launch_config": want, "summary": t["summary"]}
            if "valu_per_pair" in e:
                # instruction issue of the same launch shape (SQ_INSTS_VALU,
                # GRBM_GUI_ACTIVE): against the 2 vector instructions per CU
                # per cycle the published FP32 peak implies (1 for FP64)
                peak = 2.0 if precision == "fp32c" else 1.0
                out["issue"] = {"valu_per_pair": e["valu_per_pair"], "issue_per_cu_cycle": e["issue_per_cu_cycle"],
                                "clock_ghz": e["clock_ghz"], "issue_peak_per_cu_cycle": peak,
                                "issue_frac": e["issue_per_cu_cycle"] / peak}
            return out
    return None


def fp32c_error(system):
    """Normwise error of one FP32C right-hand side against FP64 on the same
    state: max_i |x_i - x_ref,i| / max_i |x_ref,i| and the 2-norm ratio."""
    from pawan_amd import capi

    with capi.Solver(system) as g:
        u64, w64 = g.solve("fp64")
        u32, w32 = g.solve("fp32c")

    def nw(a, b):
        return float(np.max(np.linalg.norm(a - b, axis=1)) / np.max(np.linalg.norm(b, axis=1)))

    return {"velocity_maxnorm": nw(u32, u64), "stretching_maxnorm": nw(w32, w64),
            "velocity_2norm": float(np.linalg.norm(u32 - u64) / np.linalg.norm(u64)),
            "stretching_2norm": float(np.linalg.norm(w32 - w64) / np.linalg.norm(w64))}


def _spread(n: int, count: int) -> np.ndarray:
    """count indices spread over [0, n), in an order whose every prefix
    covers the range evenly."""
    idx = np.linspace(0, n - 1, count).astype(np.int64)
    order = np.argsort((np.arange(count) * 2654435761) % count)
    return idx[order]


def cpu_baseline(system, seconds: float):
    """The oracle on the host cores, two ways, on samples of the same system.

    (a) the reference's OpenMP loop structure (parallel.cpp:14-49): serial
        source loop, fork/join parallel-for over the remaining targets of each
        source row, symmetric pairs;
    (b) the same pair arithmetic parallel over targets in one parallel region
        (each ordered pair evaluated once from the target's side).
    One right-hand side is extrapolated from each measured rate; the faster
    of the two is reported (the conservative baseline).
    """
    from oracle import oracle as O

    n = system.n
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    half = seconds / 2.0
    # (a) source rows
    rows = _spread(n, 4096)
    done, sym_pairs, ta = 0, 0, 0.0
    while done < len(rows) and ta < half:
        r = rows[done:done + 4]
        t = time.perf_counter()
        O.solve_omp_rows(system, r)
        ta += time.perf_counter() - t
        sym_pairs += int(np.sum(n - 1 - r))
        done += len(r)
    t_rhs_a = (n * (n - 1) / 2) / (sym_pairs / ta)
    # (b) targets
    tg = _spread(n, 65536)
    k, ord_pairs, tb = 0, 0, 0.0
    step = 8 * threads
    while k < len(tg) and tb < half:
        t = time.perf_counter()
        O.rates_targets(system, tg[k:k + step])
        tb += time.perf_counter() - t
        ord_pairs += (n - 1) * len(tg[k:k + step])
        k += step
    t_rhs_b = (n * (n - 1)) / (ord_pairs / tb)
    best = min(t_rhs_a, t_rhs_b)
    return {
        "value": 1.0 / (4.0 * best),
        "unit": "RK4 steps/s",
        "cores": threads,
        "kind": "port",
        "interactions_per_s": n * (n - 1) / best,
        "reference_loops_steps_per_s": 1.0 / (4.0 * t_rhs_a),
        "target_parallel_steps_per_s": 1.0 / (4.0 * t_rhs_b),
        "sample": (f"N={n} system, {threads} OpenMP threads: (a) {done} source rows through the "
                   f"restatement of parallel.cpp:14-49 ({sym_pairs:.3e} symmetric pairs, {ta:.1f} s); "
                   f"(b) {k} targets against all sources in one parallel region "
                   f"({ord_pairs:.3e} ordered pairs, {tb:.1f} s); one RHS extrapolated from each "
                   f"rate, the faster reported ({'a' if t_rhs_a <= t_rhs_b else 'b'}); "
                   f"one RK4 step = 4 RHS"),
    }


def trajectory(steps: int, every: int, n: int, tune: bool) -> dict:
    """FP32C and FP64 RK4 trajectories of the helical wake (N = n, dt =
    1e-3) side by side, graph-replayed with autotuned launches.  Every
    `every` steps: the FP32C state against the FP64 state (positions as the
    displacement since t = 0, max-norm; strengths max-norm), the FP32C
    right-hand side against an FP64 one on the FP32C state, and the
    conserved quantities of both."""
    from pawan_amd import capi, workloads

    system, dt = workloads.helix_wake(n), 1e-3
    out = {"workload": f"helix_{n}", "steps": steps, "dt": dt, "every": every, "t": [],
           "displacement_maxnorm": [], "alpha_maxnorm": [], "rhs_velocity_maxnorm": [],
           "rhs_stretching_maxnorm": [], "diag_fp64": [], "diag_fp32c": []}
    t0 = time.perf_counter()
    with capi.Solver(system) as g64, capi.Solver(system) as g32:
        if tune:
            g64.autotune("fp64")
            g32.autotune("fp32c")
        g64.set_graph(True)
        g32.set_graph(True)
        for k in range(0, steps + 1, every):
            if k > 0:
                for j in range(every):
                    g64.step("rk4", dt, "fp64", 1)
                    g32.step("rk4", dt, "fp32c", 1)
                    g32.synchronize()
                    log(f"trajectory step {k - every + j + 1}/{steps}: {time.perf_counter() - t0:.1f} s")
            p64, a64 = g64.state()
            p32, a32 = g32.state()
            out["t"].append(k * dt)
            if k > 0:
                out["displacement_maxnorm"].append(maxnorm(p32 - system.pos, p64 - system.pos))
            else:
                out["displacement_maxnorm"].append(0.0)
            out["alpha_maxnorm"].append(maxnorm(a32, a64))
            u32, w32 = g32.solve("fp32c")
            u64, w64 = g32.solve("fp64")
            out["rhs_velocity_maxnorm"].append(maxnorm(u32, u64))
            out["rhs_stretching_maxnorm"].append(maxnorm(w32, w64))
            out["diag_fp64"].append(g64.diagnose()[:15].tolist())
            out["diag_fp32c"].append(g32.diagnose()[:15].tolist())
            log(f"trajectory t={k * dt:.3f}: displacement {out['displacement_maxnorm'][-1]:.2e} "
                f"alpha {out['alpha_maxnorm'][-1]:.2e} rhs u {out['rhs_velocity_maxnorm'][-1]:.2e}")
    out["seconds"] = time.perf_counter() - t0
    return out


def device_copy(nbytes: int) -> None:
    """One device-to-device copy of nbytes through the HIP runtime (its blit
    kernel, __amd_rocclr_copyBuffer, reads and writes nbytes), after a fill
    of the source: the known-size copy that counters in KiB per dispatch are
    scaled by."""
    import ctypes

    hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so.7")
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    n = ctypes.c_size_t(nbytes)

    def ok(rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed: {rc}")

    ok(hip.hipMalloc(ctypes.byref(src), n), "hipMalloc")
    try:
        ok(hip.hipMalloc(ctypes.byref(dst), n), "hipMalloc")
        try:
            ok(hip.hipMemset(src, 1, n), "hipMemset")
            ok(hip.hipMemcpy(dst, src, n, 3), "hipMemcpy")  # 3: hipMemcpyDeviceToDevice
            ok(hip.hipDeviceSynchronize(), "hipDeviceSynchronize")
        finally:
            hip.hipFree(dst)
    finally:
        hip.hipFree(src)


def profile_rhs(args) -> dict:
    """Right-hand sides of one workload (for rocprofv3 counter and trace
    runs), after a plain device copy of known size (--copy-mb MiB): the copy
    reads and writes a known number of bytes, so counters in KiB per
    dispatch scale to bytes as factor = counter(copy) / (copy bytes / 1024)."""
    from pawan_amd import capi, workloads

    nbytes = args.copy_mb << 20
    device_copy(nbytes)
    log(f"copy of {nbytes} bytes done")
    system = workloads.helix_wake(args.n) if args.workload == "helix" else workloads.WORKLOADS[args.workload]["make"]()
    rec = {"mode": "profile-rhs", "workload": system.name, "particles": system.n, "precision": args.precision,
           "copy_bytes": nbytes, "rhs_s": []}
    with capi.Solver(system) as g:
        if args.kernel == "highorder":
            g.set_kernel(capi.PW_KERNEL_HIGHORDER)
        if args.precision == "learned":
            if args.weights:
                w = np.load(args.weights, allow_pickle=False)
                g.ff_set_weights(w["weights"])
                g.ff_set_err_coef(float(w["coef_md"]), float(w["coef_mlp"]))
            g.ff_set_params(args.theta, args.tol, use_mlp=bool(args.weights))
        elif args.config:
            v = [int(x) for x in args.config.split(",")] + [0, 0]
            g.set_pair_config(args.precision, v[0], bool(v[1]), bool(v[2]), bool(v[3]), bool(v[4]))
        g.set_profiling(True)
        for k in range(args.reps):
            t = time.perf_counter()
            g.solve(args.precision)
            rec["rhs_s"].append(time.perf_counter() - t)
            log(f"{system.name} N={system.n} {args.precision} rhs {k + 1}/{args.reps}: {rec['rhs_s'][-1]:.3f} s")
        ms, launches = g.pair_timing()
        rec["kernel_ms_avg"] = ms / max(1, launches)
        if args.precision == "learned":
            rec["tree"] = g.ff_info()
            rec["stages_ms"] = g.ff_timing()
        g.set_profiling(False)
    return rec


def profile_vring() -> dict:
    """Strength resolution of two rings, three resolves each: the
    reference's own (R = 1, core 0.1, 3 layers, 49 stations, sigma 0.1924;
    49 cells, 2,401 particles) and the largest the entry point takes (15
    layers, 961 cells, 256 stations: 246,016 particles)."""
    from pawan_amd import capi

    rec = {"mode": "profile-vring", "rings": []}
    for nl, nazi, sig in ((3, 49, 0.1924), (15, 256, 0.02)):
        pos, rad, vol = capi.vring_particles(1.0, 0.1, nl, nazi, sig)
        times = []
        with capi.Solver() as g:
            g.add_wake(pos, np.zeros_like(pos), rad, vol)
            for k in range(3):
                t = time.perf_counter()
                g.vring_resolve(0, nl, nazi, 1.0, 0.1)
                times.append(time.perf_counter() - t)
                log(f"vring L={nl} nAzi={nazi} N={len(pos)} resolve {k + 1}/3: {times[-1]:.4f} s")
        rec["rings"].append({"layers": nl, "stations": nazi, "particles": len(pos), "resolve_s": times})
    return rec


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1 << 20, help="helix particles (default 1,048,576)")
    ap.add_argument("--workload", default="helix",
                    choices=["helix", "batched_256", "leapfrog_64k", "ring_1k"])
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--fp64-steps", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--learned-steps", type=int, default=3)
    ap.add_argument("--learned-err", type=float, default=1e-4)
    ap.add_argument("--theta", type=float, default=0.5)
    ap.add_argument("--train-n", type=int, default=0,
                    help="particles of the training wakes (default: the benchmark's N)")
    ap.add_argument("--train-samples", type=int, default=2_400_000)
    ap.add_argument("--train-steps", type=int, default=40000)
    ap.add_argument("--no-drift", action="store_true")
    ap.add_argument("--drift-steps", type=int, default=200)
    ap.add_argument("--drift-every", type=int, default=10)
    ap.add_argument("--save-weights", default="", help="write the trained network (npz) here")
    # other measurements
    ap.add_argument("--trajectory", action="store_true", help="FP32C vs FP64 trajectory of the helical wake")
    ap.add_argument("--trajectory-steps", type=int, default=50)
    ap.add_argument("--profile-rhs", action="store_true", help="right-hand sides + a copy of known size")
    ap.add_argument("--profile-vring", action="store_true", help="ring strength resolutions")
    ap.add_argument("--probe-valu", action="store_true", help="vector-issue ceilings of the device (pw_valu_ceiling)")
    ap.add_argument("--precision", default="fp32c", choices=["fp64", "fp64_ref", "fp32c", "learned"])
    ap.add_argument("--reps", type=int, default=1)
    ap.add_argument("--copy-mb", type=int, default=256)
    ap.add_argument("--tol", type=float, default=0.0,
                    help="profile-rhs, tree path: absolute far-field tolerance (0: opening angle only)")
    ap.add_argument("--weights", default="", help="profile-rhs, tree path: network (npz from --save-weights)")
    ap.add_argument("--config", default="",
                    help="profile-rhs: pair-kernel launch shape 'parts,persistent,waves8[,scalar_feed[,two_per_lane]]' "
                         "(e.g. 8,1,1 or 4,1,1,1)")
    ap.add_argument("--kernel", default="gaussian", choices=["gaussian", "highorder"])
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is design-only "
                                   "(target sharding with an all-gather of sources per RK stage)",
                          "n_gpus": args.gpus}))
        return 2

    from pawan_amd import capi, workloads

    if capi.device_count() < 1:
        log("no HIP device visible")
        return 1
    if args.profile_rhs:
        print(json.dumps(profile_rhs(args)), flush=True)
        return 0
    if args.profile_vring:
        print(json.dumps(profile_vring()), flush=True)
        return 0
    if args.probe_valu:
        print(json.dumps({"mode": "probe-valu", **valu_ceilings()}), flush=True)
        return 0
    if args.trajectory:
        print(json.dumps({"mode": "trajectory", **trajectory(args.trajectory_steps, 10, args.n, not args.no_tune)}),
              flush=True)
        return 0
    t = time.perf_counter()
    if args.workload == "helix":
        system, dt = workloads.helix_wake(args.n), 1e-3
        wdesc = (f"helical rotor wake N={system.n} (4 tip helices + inboard sheet), "
                 f"RK4 dt=1e-3, FP32-compensated pairs, FP64 state")
    else:
        spec = workloads.WORKLOADS[args.workload]
        system, dt = spec["make"](), spec["dt"]
        wdesc = (f"{args.workload}: {system.nw} {'coupled' if system.coupled else 'independent'} "
                 f"wake(s), N={system.n}, RK4 dt={dt}, FP32-compensated pairs, FP64 state")
    log(f"generated {system.name}: N={system.n} in {time.perf_counter() - t:.2f} s")
    pairs = system.ordered_pairs()

    wall32, ms32, l32, tune32 = time_path(system, "fp32c", dt, args.steps, args.warmup,
                                          not args.no_tune)
    sps32 = args.steps / wall32
    avg32 = ms32 / max(1, l32)
    log(f"fp32c: {sps32:.4f} steps/s, pair kernel avg {avg32:.1f} ms over {l32} launches")

    ceil = valu_ceilings()
    log(f"valu ceilings: {ceil}")
    out64 = {}
    if not args.no_fp64 and args.fp64_steps > 0:
        wall64, ms64, l64, tune64 = time_path(system, "fp64", dt, args.fp64_steps, 0,
                                              not args.no_tune)
        avg64 = ms64 / max(1, l64)
        ach64 = FLOP_PER_PAIR * pairs / (avg64 * 1e-3) / 1e12
        out64 = {
            "fp64_launch_config": tune64,
            "fp64_steps_per_s": args.fp64_steps / wall64,
            "fp64_pair_kernel_ms": avg64,
            "fp64_tflops": ach64,
            "fp64_frac_of_peak": ach64 / PEAK_FP64_TFLOPS,
            "roofline_fp64": {
                "bound": "valu", "kernel": pair_kernel_name("fp64", tune64), "achieved": ach64, "peak": PEAK_FP64_TFLOPS,
                "unit": "TFLOP/s", "frac": ach64 / PEAK_FP64_TFLOPS,
                "traffic": (lambda t: t["bytes"] if t else None)(
                    measured_traffic(args.workload, system.n, "fp64", tune64)),
                "traffic_detail": measured_traffic(args.workload, system.n, "fp64", tune64),
                "issue_measured": measured_issue(measured_traffic(args.workload, system.n, "fp64", tune64), pairs,
                                                 avg64, ceil, "fp64_fma"),
                "avg_launch_ms": avg64, "flop_per_launch": FLOP_PER_PAIR * pairs, "launch_config": tune64},
        }
        log(f"fp64: {out64['fp64_steps_per_s']:.4f} steps/s, pair kernel avg {avg64:.1f} ms")
        out64["fp32c_error_vs_fp64"] = fp32c_error(system)
        log(f"fp32c error vs fp64: {out64['fp32c_error_vs_fp64']}")

    achieved = FLOP_PER_PAIR * pairs / (avg32 * 1e-3) / 1e12
    traffic = measured_traffic(args.workload, system.n, "fp32c", tune32)
    result = {
        "metric": METRIC,
        "value": sps32,
        "unit": "RK4 steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall32 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": wdesc,
            "particles": system.n,
            "wakes": system.nw,
            "ordered_pairs_per_rhs": pairs,
            "parallelism": "single GPU",
        },
        "interactions_per_s": 4 * pairs * sps32,
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": traffic["bytes"] if traffic else None,
            "traffic_detail": traffic,
            # achieved counts the convention's 88 FLOP per pair; the kernel
            # executes fewer (the separated-range sums), so frac can pass 1:
            # traffic_detail.issue has the executed instructions and their
            # issue rate
            "flop_convention": "88 FLOP per ordered pair",
            # the same launches' vector issue against the rate this device
            # reaches for FP32 FMAs on vector registers
            "issue_measured": measured_issue(traffic, pairs, avg32, ceil, "fp32_fma"),
            "valu_ceilings": ceil,
            "algorithmic_bytes_per_launch": 112 * system.n,
            "kernel": pair_kernel_name("fp32c", tune32),
            "launch_config": tune32,
            "avg_launch_ms": avg32,
            "flop_per_launch": FLOP_PER_PAIR * pairs,
        },
        "learned_steps_per_s": None,
    }
    result.update(out64)
    if out64:
        result["fp32c_over_fp64"] = sps32 / out64["fp64_steps_per_s"]
    if not args.no_learned and system.coupled and system.nw == 1:
        lg = learned_leg(system, dt, args.learned_steps, args.learned_err, args.theta, args.train_n or system.n,
                         args.train_samples, args.train_steps)
        result["learned"] = lg
        best = max((v for k, v in lg.items() if k in ("analytic", "mlp") and v), default=None,
                   key=lambda v: v["steps_per_s"])
        if lg.get("mlp"):
            result["learned_steps_per_s"] = lg["mlp"]["steps_per_s"]
            result["learned_over_fp32c"] = lg["mlp"]["steps_per_s"] / sps32
        if best is not None:
            result["learned_best_steps_per_s"] = best["steps_per_s"]
        cfg = lg.pop("_drift_cfg", None)
        if args.save_weights and cfg is not None:
            np.savez(args.save_weights, weights=cfg["weights"], coef_md=cfg["coef_md"], coef_mlp=cfg["coef_mlp"])
        if not args.no_drift:
            result["drift"] = drift_leg(args.drift_steps, args.drift_every, cfg, args.learned_err)
    if not args.no_cpu and system.coupled:
        log("cpu baseline ...")
        result["cpu_baseline"] = cpu_baseline(system, args.cpu_seconds)
        log(f"cpu baseline: {result['cpu_baseline']['value']:.3e} steps/s on "
            f"{result['cpu_baseline']['cores']} threads")
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
