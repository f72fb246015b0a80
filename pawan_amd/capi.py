"""ctypes binding of the C-ABI in include/pawan_gpu.h (libpawan_gpu.so).

Thin Python access to the GPU solver for tests and benchmarks.  There is no
CPU fallback: if the library is missing or no HIP device is visible, calls
raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import workloads as _wl

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libpawan_gpu.so")

PW_OK, PW_ERR_ARG, PW_ERR_HIP, PW_ERR_STATE, PW_ERR_NONFINITE, PW_ERR_UNSUPPORTED, PW_ERR_NOMEM = range(7)
PW_FP64, PW_FP32C, PW_LEARNED, PW_FP64_REF = 0, 1, 2, 3
PW_EULER, PW_RK3, PW_RK4 = 1, 3, 4
PW_KERNEL_GAUSSIAN, PW_KERNEL_HIGHORDER = 0, 1
PW_VARIANT_WAVES8, PW_VARIANT_SCALAR_FEED, PW_VARIANT_TWO_PER_LANE = 1, 2, 4

PRECISIONS = {"fp64": PW_FP64, "fp64_ref": PW_FP64_REF, "fp32c": PW_FP32C, "learned": PW_LEARNED}
SCHEMES = {"euler": PW_EULER, "rk3": PW_RK3, "rk4": PW_RK4}

# Every extern "C" symbol declared in include/pawan_gpu.h.
EXPORTS = [
    "pw_version", "pw_last_error", "pw_device_count", "pw_create", "pw_destroy",
    "pw_add_wake", "pw_set_nu", "pw_get_nu", "pw_set_coupled", "pw_get_size",
    "pw_set_wake_state", "pw_get_wake_state", "pw_get_wake_rates", "pw_set_states",
    "pw_get_states", "pw_get_rates", "pw_solve", "pw_step", "pw_diagnose",
    "pw_synchronize", "pw_set_profiling", "pw_get_pair_timing",
    "pw_set_pair_config", "pw_get_pair_config", "pw_autotune", "pw_set_graph",
    "pw_ff_set_params", "pw_ff_get_params", "pw_ff_set_err_coef", "pw_ff_get_err_coef",
    "pw_ff_build", "pw_ff_info", "pw_ff_export", "pw_ff_sample", "pw_ff_clear_samples",
    "pw_ff_init_weights", "pw_ff_set_weights", "pw_ff_get_weights", "pw_ff_train",
    "pw_ff_mlp_forward", "pw_ff_mlp_grad", "pw_ff_set_conserve", "pw_ff_get_conserve",
    "pw_velocity_at", "pw_set_kernel", "pw_get_kernel", "pw_vring_particles", "pw_vring_resolve",
    "pw_set_compensation", "pw_ff_get_timing", "pw_interact_wakes", "pw_vorticity_field",
    "pw_diagnose_wakes", "pw_velocity_at_noalloc", "pw_wake_split", "pw_wake_merge", "pw_set_wake_volume",
    "pw_valu_ceiling",
]

# pw_ff_export selectors
(FF_PERM, FF_NODES, FF_NEAR_CNT, FF_NEAR, FF_FAR_CNT, FF_FAR, FF_SAMPLES_X, FF_SAMPLES_Y, FF_SAMPLES_AUX,
 FF_QUAD) = range(10)
MLP_PARAMS = 36748
FF_INFO_NAMES = ["n", "leaves", "nodes", "levels", "cap_near", "cap_far", "samples", "near_pairs",
                 "far_evals", "overflow", "max_near", "max_far", "near_core_pairs", "mlp_evals"]
FF_STAGE_NAMES = ["tree_lists_ms", "near_ms", "far_md_ms", "far_mlp_ms", "rest_ms"]
FF_TRAIN_NAMES = ["train_loss", "val_loss", "val_loss_zero", "rms_ratio_u", "rms_ratio_grad",
                  "coef_md", "coef_mlp", "steps", "val_samples", "ms"]


class PawanError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_LIB: C.CDLL | None = None


def load(path: str | None = None) -> C.CDLL:
    """Load libpawan_gpu.so (raises if it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"{path} is missing: build it with `python -m pawan_amd.build` "
            "(or __graft_entry__.build()); there is no CPU fallback")
    lib = C.CDLL(path)
    vp, i64, d, i = C.c_void_p, C.c_int64, C.c_double, C.c_int
    pvp = C.POINTER(C.c_void_p)
    pi64 = C.POINTER(C.c_int64)
    sig = {
        "pw_version": ([], C.c_char_p),
        "pw_last_error": ([], C.c_char_p),
        "pw_device_count": ([C.POINTER(C.c_int)], i),
        "pw_create": ([pvp, i], i),
        "pw_destroy": ([vp], i),
        "pw_add_wake": ([vp, i64, i64, _dp, _dp, _dp, _dp, C.POINTER(C.c_int)], i),
        "pw_set_nu": ([vp, d], i),
        "pw_get_nu": ([vp, C.POINTER(C.c_double)], i),
        "pw_set_coupled": ([vp, i], i),
        "pw_set_kernel": ([vp, i], i),
        "pw_set_compensation": ([vp, i], i),
        "pw_vring_particles": ([d, d, i, i, d, _dp, _dp, _dp], i),
        "pw_vring_resolve": ([vp, i, i, i, d, d, C.c_void_p, C.c_void_p], i),
        "pw_get_kernel": ([vp, C.POINTER(C.c_int)], i),
        "pw_get_size": ([vp, C.POINTER(C.c_int64), C.POINTER(C.c_int)], i),
        "pw_set_wake_state": ([vp, i, _dp, _dp], i),
        "pw_set_wake_volume": ([vp, i, _dp], i),
        "pw_get_wake_state": ([vp, i, _dp, _dp], i),
        "pw_get_wake_rates": ([vp, i, _dp, _dp], i),
        "pw_set_states": ([vp, _dp, i64], i),
        "pw_get_states": ([vp, _dp, i64], i),
        "pw_get_rates": ([vp, _dp, i64], i),
        "pw_solve": ([vp, i], i),
        "pw_velocity_at": ([vp, _dp, i64, _dp], i),
        "pw_velocity_at_noalloc": ([vp, _dp, i64, _dp], i),
        "pw_wake_split": ([i, d, i64, pi64, pi64, _dp, _dp, _dp, _dp, _dp, _dp, pi64], i),
        "pw_wake_merge": ([i, d, i64, _dp, _dp, _dp, _dp, _dp, pi64], i),
        "pw_step": ([vp, i, d, i, i64], i),
        "pw_diagnose": ([vp, _dp, i64], i),
        "pw_synchronize": ([vp], i),
        "pw_set_profiling": ([vp, i], i),
        "pw_get_pair_timing": ([vp, C.POINTER(C.c_double), C.POINTER(C.c_int64)], i),
        "pw_set_pair_config": ([vp, i, i, i, i], i),
        "pw_get_pair_config": ([vp, i, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)], i),
        "pw_autotune": ([vp, i, C.POINTER(C.c_double)], i),
        "pw_set_graph": ([vp, i], i),
        "pw_ff_set_params": ([vp, d, d, i], i),
        "pw_ff_get_params": ([vp, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int)], i),
        "pw_ff_set_err_coef": ([vp, d, d], i),
        "pw_ff_set_conserve": ([vp, i], i),
        "pw_ff_get_conserve": ([vp, C.POINTER(C.c_int)], i),
        "pw_ff_get_err_coef": ([vp, C.POINTER(C.c_double), C.POINTER(C.c_double)], i),
        "pw_ff_build": ([vp], i),
        "pw_ff_info": ([vp, C.POINTER(C.c_int64), i64], i),
        "pw_ff_get_timing": ([vp, C.POINTER(C.c_double), i64], i),
        "pw_interact_wakes": ([vp, i, C.c_void_p, i], i),
        "pw_vorticity_field": ([vp, C.c_void_p, i, _dp, i64], i),
        "pw_diagnose_wakes": ([vp, i, i, _dp, i64], i),
        "pw_ff_export": ([vp, i, vp, i64], i),
        "pw_ff_sample": ([vp, i64, C.c_uint64], i),
        "pw_ff_clear_samples": ([vp], i),
        "pw_ff_init_weights": ([vp, C.c_uint64], i),
        "pw_ff_set_weights": ([vp, _fp, i64], i),
        "pw_ff_get_weights": ([vp, _fp, i64], i),
        "pw_ff_train": ([vp, i64, i, d, C.c_uint64, _dp, i64], i),
        "pw_ff_mlp_forward": ([vp, _fp, i64, _fp], i),
        "pw_ff_mlp_grad": ([vp, _fp, _fp, i64, _fp, C.POINTER(C.c_double)], i),
        "pw_valu_ceiling": ([i, i, C.POINTER(C.c_double), C.POINTER(C.c_int)], i),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _LIB = lib
    return lib


def _check(rc: int) -> None:
    if rc != PW_OK:
        raise PawanError(rc, load().pw_last_error().decode())


def device_count() -> int:
    n = C.c_int(0)
    rc = load().pw_device_count(C.byref(n))
    return n.value if rc == PW_OK else 0


VALU_FORMS = {"fp32_fma": 0, "fp32_fma_sgpr": 1, "fp64_fma": 2, "fp32_mul": 3}


def valu_ceiling(form: str, waves_per_cu: int = 32) -> tuple[float, int]:
    """Vector-issue rate the device reaches for one instruction form (wave
    instructions per CU per nanosecond; forms in VALU_FORMS) and the
    device's CU count."""
    r, cus = C.c_double(0.0), C.c_int(0)
    _check(load().pw_valu_ceiling(VALU_FORMS[form], int(waves_per_cu), C.byref(r), C.byref(cus)))
    return r.value, cus.value


def vring_particles(R: float, core: float, nlayers: int, nazi: int, sig: float):
    """Particles of __vring(R, core, nlayers, nazi, sig) (vring.cpp:50-100):
    (pos, radius, volume), vorticity zero until resolved."""
    n = nazi * (2 * nlayers + 1) ** 2
    pos, radius, volume = np.zeros((n, 3)), np.zeros(n), np.zeros(n)
    _check(load().pw_vring_particles(float(R), float(core), int(nlayers), int(nazi), float(sig),
                                     pos, radius, volume))
    return pos, radius, volume


class Solver:
    """One pw_ctx: a set of wakes on one GPU (the reference's __interaction)."""

    def __init__(self, system: "_wl.ParticleSystem | None" = None, device: int = 0,
                 nu: float | None = None, coupled: bool | None = None):
        self._lib = load()
        h = C.c_void_p()
        _check(self._lib.pw_create(C.byref(h), device))
        self._h = h
        self.wake_sizes: list[int] = []
        self.wake_max: list[int] = []
        if system is not None:
            if coupled is None:
                coupled = system.coupled
            self.set_coupled(coupled)
            for w in range(system.nw):
                a, b = int(system.off[w]), int(system.off[w + 1])
                self.add_wake(system.pos[a:b], system.alpha[a:b], system.radius[a:b],
                              system.volume[a:b])
            self.set_nu(system.nu if nu is None else nu)
        elif nu is not None:
            self.set_nu(nu)

    # -- lifetime ----------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.pw_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- configuration -------------------------------------------------------
    def add_wake(self, pos, alpha, radius, volume, maxn: int | None = None) -> int:
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        n = len(pos)
        alpha = np.ascontiguousarray(alpha, dtype=np.float64).reshape(n, 3)
        radius = np.ascontiguousarray(radius, dtype=np.float64).reshape(n)
        volume = np.ascontiguousarray(volume, dtype=np.float64).reshape(n)
        maxn = n if maxn is None else int(maxn)
        w = C.c_int(-1)
        _check(self._lib.pw_add_wake(self._h, n, maxn, pos, alpha, radius, volume, C.byref(w)))
        self.wake_sizes.append(n)
        self.wake_max.append(maxn)
        return w.value

    def set_nu(self, nu: float) -> None:
        _check(self._lib.pw_set_nu(self._h, float(nu)))

    @property
    def nu(self) -> float:
        v = C.c_double()
        _check(self._lib.pw_get_nu(self._h, C.byref(v)))
        return v.value

    def set_coupled(self, coupled: bool) -> None:
        _check(self._lib.pw_set_coupled(self._h, int(bool(coupled))))

    def vring_resolve(self, wake: int, nlayers: int, nazi: int, R: float, core: float):
        """__resolve::rebuild on the ring held by `wake`: returns the cell
        strengths and the influence matrix; the wake's vorticity is set."""
        m = (2 * nlayers + 1) ** 2
        s = np.zeros(m)
        M = np.zeros((m, m))
        _check(self._lib.pw_vring_resolve(self._h, int(wake), int(nlayers), int(nazi), float(R), float(core),
                                          s.ctypes.data_as(C.c_void_p), M.ctypes.data_as(C.c_void_p)))
        return s, M

    def set_kernel(self, kernel: int) -> None:
        """PW_KERNEL_GAUSSIAN or PW_KERNEL_HIGHORDER."""
        _check(self._lib.pw_set_kernel(self._h, int(kernel)))

    def set_compensation(self, on: bool) -> None:
        """FP32C: compensated per-target sums (default) or plain FP32 sums
        (a measuring aid)."""
        _check(self._lib.pw_set_compensation(self._h, 1 if on else 0))

    @property
    def kernel(self) -> int:
        v = C.c_int()
        _check(self._lib.pw_get_kernel(self._h, C.byref(v)))
        return v.value

    @property
    def n(self) -> int:
        return int(sum(self.wake_sizes))

    @property
    def state_len(self) -> int:
        return 6 * int(sum(self.wake_max))

    # -- state ---------------------------------------------------------------
    def set_wake_state(self, w: int, pos, alpha) -> None:
        n = self.wake_sizes[w]
        _check(self._lib.pw_set_wake_state(
            self._h, w, np.ascontiguousarray(pos, dtype=np.float64).reshape(n, 3),
            np.ascontiguousarray(alpha, dtype=np.float64).reshape(n, 3)))

    def wake_state(self, w: int):
        n = self.wake_sizes[w]
        pos, alpha = np.zeros((n, 3)), np.zeros((n, 3))
        _check(self._lib.pw_get_wake_state(self._h, w, pos, alpha))
        return pos, alpha

    def wake_rates(self, w: int):
        n = self.wake_sizes[w]
        vel, ret = np.zeros((n, 3)), np.zeros((n, 3))
        _check(self._lib.pw_get_wake_rates(self._h, w, vel, ret))
        return vel, ret

    def state(self):
        """(pos, alpha) of all particles, wakes concatenated."""
        parts = [self.wake_state(w) for w in range(len(self.wake_sizes))]
        return np.concatenate([p for p, _ in parts]), np.concatenate([a for _, a in parts])

    def rates(self):
        """(vel, retvor) of all particles from the last right-hand side."""
        parts = [self.wake_rates(w) for w in range(len(self.wake_sizes))]
        return np.concatenate([p for p, _ in parts]), np.concatenate([a for _, a in parts])

    def set_states(self, state: np.ndarray) -> None:
        state = np.ascontiguousarray(state, dtype=np.float64)
        _check(self._lib.pw_set_states(self._h, state, len(state)))

    def get_states(self) -> np.ndarray:
        out = np.zeros(self.state_len)
        _check(self._lib.pw_get_states(self._h, out, len(out)))
        return out

    def get_rates(self) -> np.ndarray:
        out = np.zeros(self.state_len)
        _check(self._lib.pw_get_rates(self._h, out, len(out)))
        return out

    # -- compute -------------------------------------------------------------
    def solve(self, precision: str = "fp64"):
        _check(self._lib.pw_solve(self._h, PRECISIONS[precision]))
        return self.rates()

    @staticmethod
    def _pairs(pairs):
        if pairs is None:
            return None, 0
        arr = np.ascontiguousarray(np.asarray(pairs, dtype=np.int32).reshape(-1, 2))
        return arr, len(arr)

    def interact_wakes(self, precision: str, pairs):
        """Rates of the listed wake blocks only [(a, b), a <= b]."""
        arr, n = self._pairs(pairs)
        _check(self._lib.pw_interact_wakes(self._h, PRECISIONS[precision],
                                           arr.ctypes.data_as(C.c_void_p) if n else None, n))
        return self.rates()

    def vorticity_field(self, pairs=None) -> np.ndarray:
        """Smoothed vorticity field at every particle (N x 3) of the listed
        wake blocks, or of the whole system."""
        arr, n = self._pairs(pairs)
        out = np.zeros(3 * self.n)
        _check(self._lib.pw_vorticity_field(self._h, arr.ctypes.data_as(C.c_void_p) if arr is not None else None,
                                            n, out, len(out)))
        return out.reshape(-1, 3)

    def diagnose_wakes(self, a: int, b: int) -> np.ndarray:
        out = np.zeros(16)
        _check(self._lib.pw_diagnose_wakes(self._h, int(a), int(b), out, 16))
        return out

    def velocity_at(self, points) -> np.ndarray:
        """Velocity (FP64) at points (m x 3) from every particle of the
        context, each with its own radius (getVi, interaction.cpp:453-476)."""
        pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        u = np.empty_like(pts)
        _check(self._lib.pw_velocity_at(self._h, pts, len(pts), u))
        return u

    def velocity_at_noalloc(self, points) -> np.ndarray:
        """velocity_at through pw_velocity_at_noalloc: no device allocation,
        at most 65,536 points, the particles already on the device."""
        pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        u = np.empty_like(pts)
        _check(self._lib.pw_velocity_at_noalloc(self._h, pts, len(pts), u))
        return u

    def step(self, scheme: str, dt: float, precision: str = "fp64", nsteps: int = 1) -> None:
        _check(self._lib.pw_step(self._h, SCHEMES[scheme], float(dt), PRECISIONS[precision],
                                 int(nsteps)))

    def diagnose(self) -> np.ndarray:
        out = np.zeros(15 * max(1, len(self.wake_sizes)))
        _check(self._lib.pw_diagnose(self._h, out, len(out)))
        return out

    def synchronize(self) -> None:
        _check(self._lib.pw_synchronize(self._h))

    def set_profiling(self, on: bool) -> None:
        _check(self._lib.pw_set_profiling(self._h, int(bool(on))))

    def set_pair_config(self, precision: str, parts: int = 1, persistent: bool = False,
                        waves8: bool = False, scalar_feed: bool = False, two_per_lane: bool = False) -> None:
        variant = ((PW_VARIANT_WAVES8 if waves8 else 0) | (PW_VARIANT_SCALAR_FEED if scalar_feed else 0) |
                   (PW_VARIANT_TWO_PER_LANE if two_per_lane else 0))
        _check(self._lib.pw_set_pair_config(self._h, PRECISIONS[precision], int(parts),
                                            int(bool(persistent)), variant))

    def pair_config(self, precision: str) -> dict:
        p, q, w = C.c_int(), C.c_int(), C.c_int()
        _check(self._lib.pw_get_pair_config(self._h, PRECISIONS[precision], C.byref(p), C.byref(q),
                                            C.byref(w)))
        return {"parts": p.value, "persistent": bool(q.value), "waves8": bool(w.value & PW_VARIANT_WAVES8),
                "scalar_feed": bool(w.value & PW_VARIANT_SCALAR_FEED),
                "two_per_lane": bool(w.value & PW_VARIANT_TWO_PER_LANE)}

    def autotune(self, precision: str) -> float:
        ms = C.c_double()
        _check(self._lib.pw_autotune(self._h, PRECISIONS[precision], C.byref(ms)))
        return ms.value

    def set_graph(self, on: bool) -> None:
        _check(self._lib.pw_set_graph(self._h, int(bool(on))))

    def pair_timing(self):
        ms, n = C.c_double(), C.c_int64()
        _check(self._lib.pw_get_pair_timing(self._h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    # -- learned far field (PW_LEARNED) ---------------------------------------
    def ff_set_params(self, theta: float = 0.5, tol: float = 0.0, use_mlp: bool = True) -> None:
        _check(self._lib.pw_ff_set_params(self._h, float(theta), float(tol), int(bool(use_mlp))))

    def ff_params(self) -> dict:
        th, tol, um = C.c_double(), C.c_double(), C.c_int()
        _check(self._lib.pw_ff_get_params(self._h, C.byref(th), C.byref(tol), C.byref(um)))
        return {"theta": th.value, "tol": tol.value, "use_mlp": bool(um.value)}

    def ff_set_conserve(self, on: bool) -> None:
        _check(self._lib.pw_ff_set_conserve(self._h, int(bool(on))))

    def ff_conserve(self) -> bool:
        v = C.c_int()
        _check(self._lib.pw_ff_get_conserve(self._h, C.byref(v)))
        return bool(v.value)

    def ff_set_err_coef(self, coef_md: float, coef_mlp: float) -> None:
        _check(self._lib.pw_ff_set_err_coef(self._h, float(coef_md), float(coef_mlp)))

    def ff_err_coef(self):
        a, b = C.c_double(), C.c_double()
        _check(self._lib.pw_ff_get_err_coef(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def ff_build(self) -> None:
        _check(self._lib.pw_ff_build(self._h))

    def ff_info(self) -> dict:
        out = (C.c_int64 * len(FF_INFO_NAMES))()
        _check(self._lib.pw_ff_info(self._h, out, len(FF_INFO_NAMES)))
        return dict(zip(FF_INFO_NAMES, [int(v) for v in out]))

    def ff_timing(self) -> dict:
        """Stage times (ms, totals) of the tree-path right-hand sides since
        profiling was switched on."""
        out = (C.c_double * (1 + len(FF_STAGE_NAMES)))()
        _check(self._lib.pw_ff_get_timing(self._h, out, len(out)))
        return dict(zip(["rhs"] + FF_STAGE_NAMES, [float(v) for v in out]))

    def ff_export(self, what: int) -> np.ndarray:
        info = self.ff_info()
        nl = info["leaves"]
        shapes = {
            FF_PERM: (np.int32, (info["n"],)),
            FF_NODES: (np.float64, (info["nodes"], 24)),
            FF_NEAR_CNT: (np.int32, (nl,)),
            FF_NEAR: (np.int32, (nl, info["cap_near"])),
            FF_FAR_CNT: (np.int32, (nl,)),
            FF_FAR: (np.int32, (nl, info["cap_far"])),
            FF_SAMPLES_X: (np.float32, (info["samples"], 16)),
            FF_SAMPLES_Y: (np.float32, (info["samples"], 12)),
            FF_SAMPLES_AUX: (np.float64, (info["samples"], 12)),
            FF_QUAD: (np.float64, (info["nodes"], 20)),
        }
        dt, shape = shapes[what]
        out = np.zeros(shape, dtype=dt)
        _check(self._lib.pw_ff_export(self._h, what, out.ctypes.data_as(C.c_void_p), out.nbytes))
        return out

    def ff_sample(self, nsamples: int, seed: int = 1) -> None:
        _check(self._lib.pw_ff_sample(self._h, int(nsamples), int(seed) & (2**64 - 1)))

    def ff_clear_samples(self) -> None:
        _check(self._lib.pw_ff_clear_samples(self._h))

    def ff_init_weights(self, seed: int = 1) -> None:
        _check(self._lib.pw_ff_init_weights(self._h, int(seed) & (2**64 - 1)))

    def ff_set_weights(self, params) -> None:
        p = np.ascontiguousarray(params, dtype=np.float32).reshape(-1)
        _check(self._lib.pw_ff_set_weights(self._h, p, len(p)))

    def ff_get_weights(self) -> np.ndarray:
        out = np.zeros(MLP_PARAMS, dtype=np.float32)
        _check(self._lib.pw_ff_get_weights(self._h, out, len(out)))
        return out

    def ff_train(self, steps: int, batch: int = 1024, lr: float = 1e-3, seed: int = 1) -> dict:
        rep = np.zeros(len(FF_TRAIN_NAMES))
        _check(self._lib.pw_ff_train(self._h, int(steps), int(batch), float(lr), int(seed) & (2**64 - 1),
                                     rep, len(rep)))
        return dict(zip(FF_TRAIN_NAMES, rep.tolist()))

    def ff_mlp_forward(self, X) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, 16)
        Y = np.zeros((len(X), 12), dtype=np.float32)
        _check(self._lib.pw_ff_mlp_forward(self._h, X, len(X), Y))
        return Y

    def ff_mlp_grad(self, X, Y):
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, 16)
        Y = np.ascontiguousarray(Y, dtype=np.float32).reshape(-1, 12)
        g = np.zeros(MLP_PARAMS, dtype=np.float32)
        loss = C.c_double()
        _check(self._lib.pw_ff_mlp_grad(self._h, X, Y, len(X), g, C.byref(loss)))
        return g, loss.value


def wake_split(pos, alpha, radius, volume, birth, active, n: int, stepnum: float, maxn: int | None = None,
               device: int = 0):
    """__wake::split on the device (pw_wake_split), in place on arrays of
    maxn rows whose first n rows are particles; returns (n after, splits)."""
    maxn = len(radius) if maxn is None else maxn
    arrs = [pos, alpha, radius, volume, birth, active]
    for a in arrs:
        if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]):
            raise TypeError("split arrays must be C-contiguous float64 numpy arrays (updated in place)")
    if len(pos) < maxn or len(alpha) < maxn or min(len(radius), len(volume), len(birth), len(active)) < maxn:
        raise ValueError("arrays need maxn rows")
    nn, k, ns = C.c_int64(n), C.c_int64(n), C.c_int64(0)
    _check(load().pw_wake_split(device, float(stepnum), int(maxn), C.byref(nn), C.byref(k), pos, alpha, radius,
                                volume, birth, active, C.byref(ns)))
    return nn.value, ns.value


def wake_merge(pos, alpha, radius, volume, active, n: int, stepnum: float, device: int = 0) -> int:
    """__wake::merge on the device (pw_wake_merge), in place; returns the
    number of merges."""
    for a in (pos, alpha, radius, volume, active):
        if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]):
            raise TypeError("merge arrays must be C-contiguous float64 numpy arrays (updated in place)")
    if len(pos) < n or len(alpha) < n or min(len(radius), len(volume), len(active)) < n:
        raise ValueError("arrays need n rows")
    nm = C.c_int64(0)
    _check(load().pw_wake_merge(device, float(stepnum), int(n), pos, alpha, radius, volume, active, C.byref(nm)))
    return nm.value
