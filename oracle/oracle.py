"""ctypes binding of the oracle library (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The product package never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
RHS_FN = C.CFUNCTYPE(None, C.c_int64, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_void_p)


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _load() -> C.CDLL:
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    d, i64, dp, ip = C.c_double, C.c_int64, _dp, _ip
    pd = C.POINTER(C.c_double)
    lib.or_dnrm2.argtypes = [dp]
    lib.or_dnrm2.restype = d
    lib.or_kernel.argtypes = [d, d, pd, pd, pd, pd]
    lib.or_kernel_q128.argtypes = [d, d, pd, pd, pd]
    lib.or_qsig.argtypes = [d, d]
    lib.or_qsig.restype = d
    lib.or_interact.argtypes = [d, d, d, dp, dp, dp, dp, d, d, dp, dp, dp, dp]
    sysargs = [C.c_int, ip, dp, dp, dp, dp, d]
    lib.or_solve.argtypes = sysargs + [dp, dp]
    lib.or_solve_omp.argtypes = sysargs + [dp, dp]
    lib.or_solve_q128.argtypes = sysargs + [dp, dp]
    lib.or_solve_omp_rows.argtypes = [i64, dp, dp, dp, dp, d, ip, i64, dp, dp]
    lib.or_rates_targets.argtypes = sysargs + [ip, i64, dp, dp, C.c_int]
    lib.or_rates_targets_q128.argtypes = sysargs + [ip, i64, dp, dp]
    lib.or_step.argtypes = [C.c_int, d, i64, C.c_int, ip, dp, dp, dp, dp, d, C.c_int]
    lib.or_step.restype = C.c_int
    lib.or_step_generic.argtypes = [C.c_int, d, i64, i64, dp, RHS_FN, C.c_void_p]
    lib.or_step_generic.restype = C.c_int
    lib.or_diagnose.argtypes = [C.c_int, ip, dp, dp, dp, d, dp, C.c_int]
    lib.or_velocity_at.argtypes = [i64, dp, dp, dp, i64, dp, dp]
    lib.or_vring.argtypes = [d, d, C.c_int, C.c_int, d, dp, dp, dp, dp]
    lib.or_vring_ideal.argtypes = [d, d, C.c_int, dp]
    lib.or_vring_influence.argtypes = [C.c_int, C.c_int, dp, dp, dp]
    lib.or_lu_solve.argtypes = [C.c_int, dp, dp, dp]
    lib.or_lu_solve.restype = C.c_int
    lib.or_set_smoothing.argtypes = [C.c_int]
    lib.or_get_smoothing.restype = C.c_int
    lib.or_enst_pair.argtypes = [d, d, pd, pd]
    lib.or_enstf_pair.argtypes = [d, d]
    lib.or_enstf_pair.restype = d
    lib.or_enst_self.argtypes = [d]
    lib.or_enst_self.restype = d
    lib.or_enstf_self.argtypes = [d]
    lib.or_enstf_self.restype = d
    pi64 = C.POINTER(C.c_int64)
    lib.or_wake_split.argtypes = [d, i64, pi64, pi64, dp, dp, dp, dp, dp, dp]
    lib.or_wake_split.restype = i64
    lib.or_wake_merge.argtypes = [d, i64, dp, dp, dp, dp, dp]
    lib.or_wake_merge.restype = i64
    return lib


_LIB: C.CDLL | None = None


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        _LIB = _load()
    return _LIB


def _arrs(s):
    return (s.nw, np.ascontiguousarray(s.off, dtype=np.int64), np.ascontiguousarray(s.pos),
            np.ascontiguousarray(s.alpha), np.ascontiguousarray(s.radius),
            np.ascontiguousarray(s.volume))


def kernel(rho: float, sigma: float):
    q, F, Z, n = C.c_double(), C.c_double(), C.c_double(), C.c_double()
    lib().or_kernel(rho, sigma, q, F, Z, n)
    return q.value, F.value, Z.value, n.value


def kernel_q128(rho: float, sigma: float):
    q, F, Z = C.c_double(), C.c_double(), C.c_double()
    lib().or_kernel_q128(rho, sigma, q, F, Z)
    return q.value, F.value, Z.value


def qsig(rho: float, sigma: float) -> float:
    return lib().or_qsig(rho, sigma)


GAUSSIAN, HIGHORDER = 0, 1


class smoothing:
    """Context manager: every oracle function inside uses the given
    smoothing kernel (GAUSSIAN = gaussian.h, HIGHORDER = highorder.h; the
    reference picks one at compile time, interaction_utils.h:22-26)."""

    def __init__(self, kind: int):
        self.kind = kind

    def __enter__(self):
        self.prev = lib().or_get_smoothing()
        lib().or_set_smoothing(self.kind)
        return self

    def __exit__(self, *exc):
        lib().or_set_smoothing(self.prev)
        return False


def enst_pair(rho: float, sigma: float):
    F1, F2 = C.c_double(), C.c_double()
    lib().or_enst_pair(rho, sigma, F1, F2)
    return F1.value, F2.value


def enstf_pair(rho: float, sigma: float) -> float:
    return lib().or_enstf_pair(rho, sigma)


def enst_self(sigma: float) -> float:
    return lib().or_enst_self(sigma)


def enstf_self(sigma: float) -> float:
    return lib().or_enstf_self(sigma)


def dnrm2(x) -> float:
    return lib().or_dnrm2(np.ascontiguousarray(x, dtype=np.float64))


def interact(nu, s_src, s_trg, r_src, r_trg, a_src, a_trg, v_src, v_trg):
    """One symmetric pair; returns (dr_src, dr_trg, da_src, da_trg)."""
    out = [np.zeros(3) for _ in range(4)]
    f = lambda x: np.ascontiguousarray(x, dtype=np.float64)
    lib().or_interact(nu, s_src, s_trg, f(r_src), f(r_trg), f(a_src), f(a_trg), v_src, v_trg, *out)
    return tuple(out)


def _check_coupled(s):
    if not s.coupled and s.nw > 1:
        raise ValueError("the oracle solves one coupled system; split independent wakes first")


def solve(s, nu=None):
    """Serial reference loops (interaction.cpp:196-302)."""
    _check_coupled(s)
    nu = s.nu if nu is None else nu
    vel, ret = np.zeros((s.n, 3)), np.zeros((s.n, 3))
    lib().or_solve(*_arrs(s), nu, vel, ret)
    return vel, ret


def velocity_at(s, points):
    """getVi (interaction.cpp:453-476): velocity at each point from every
    particle of every wake, sigma = the particle's own radius."""
    pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
    u = np.zeros_like(pts)
    lib().or_velocity_at(s.n, np.ascontiguousarray(s.pos, dtype=np.float64),
                         np.ascontiguousarray(s.alpha, dtype=np.float64),
                         np.ascontiguousarray(s.radius, dtype=np.float64), len(pts), pts, u)
    return u


def solve_omp(s, nu=None):
    _check_coupled(s)
    nu = s.nu if nu is None else nu
    vel, ret = np.zeros((s.n, 3)), np.zeros((s.n, 3))
    lib().or_solve_omp(*_arrs(s), nu, vel, ret)
    return vel, ret


def solve_q128(s, nu=None):
    _check_coupled(s)
    nu = s.nu if nu is None else nu
    vel, ret = np.zeros((s.n, 3)), np.zeros((s.n, 3))
    lib().or_solve_q128(*_arrs(s), nu, vel, ret)
    return vel, ret


def rates_targets(s, targets=None, nu=None, use_omp=True):
    """Rates of the given targets, bitwise equal to :func:`solve` rows."""
    _check_coupled(s)
    nu = s.nu if nu is None else nu
    t = np.arange(s.n, dtype=np.int64) if targets is None else np.ascontiguousarray(targets, dtype=np.int64)
    vel, ret = np.zeros((len(t), 3)), np.zeros((len(t), 3))
    lib().or_rates_targets(*_arrs(s), nu, t, len(t), vel, ret, int(use_omp))
    return vel, ret


def rates_targets_q128(s, targets, nu=None):
    """__float128 rates of the given targets (all sources, no ordering)."""
    _check_coupled(s)
    nu = s.nu if nu is None else nu
    t = np.ascontiguousarray(targets, dtype=np.int64)
    vel, ret = np.zeros((len(t), 3)), np.zeros((len(t), 3))
    lib().or_rates_targets_q128(*_arrs(s), nu, t, len(t), vel, ret)
    return vel, ret


def solve_omp_rows(s, rows, nu=None):
    """Run the parallel.cpp source-row loop body for the listed rows only."""
    if s.nw != 1:
        raise ValueError("single wake only")
    nu = s.nu if nu is None else nu
    vel, ret = np.zeros((s.n, 3)), np.zeros((s.n, 3))
    r = np.ascontiguousarray(rows, dtype=np.int64)
    _, _, pos, alpha, rad, vol = _arrs(s)
    lib().or_solve_omp_rows(s.n, pos, alpha, rad, vol, nu, r, len(r), vel, ret)
    return vel, ret


SCHEMES = {"euler": 1, "rk3": 3, "rk4": 4}


def step(s, scheme: str, dt: float, nsteps: int = 1, nu=None, use_omp=True):
    """Returns a copy of ``s`` advanced by ``nsteps`` steps."""
    _check_coupled(s)
    nu = s.nu if nu is None else nu
    out = s.copy()
    rc = lib().or_step(SCHEMES[scheme], dt, nsteps, *_arrs(out)[:2], out.pos, out.alpha,
                       out.radius, out.volume, nu, int(use_omp))
    if rc != 0:
        raise ValueError(f"bad scheme {scheme}")
    return out


def step_generic(scheme: str, dt: float, nsteps: int, y: np.ndarray, f) -> np.ndarray:
    """The steppers over a flat state with a Python right-hand side f(y)->k."""
    y = np.ascontiguousarray(y, dtype=np.float64).copy()
    L = len(y)

    def cb(n, yp, kp, _ctx):
        yy = np.ctypeslib.as_array(yp, shape=(n,))
        kk = np.ctypeslib.as_array(kp, shape=(n,))
        kk[:] = f(yy.copy())

    fn = RHS_FN(cb)
    rc = lib().or_step_generic(SCHEMES[scheme], dt, nsteps, L, y, fn, None)
    if rc != 0:
        raise ValueError(f"bad scheme {scheme}")
    return y


DIAG_NAMES = ["Ox", "Oy", "Oz", "Ix", "Iy", "Iz", "Ax", "Ay", "Az", "H", "E", "EF", "KE", "KEF", "Zc"]


def diagnose(s, zc_prev: float = 0.0, use_omp=True) -> np.ndarray:
    _check_coupled(s)
    out = np.zeros(15)
    nw, off, pos, alpha, rad, _ = _arrs(s)
    lib().or_diagnose(nw, off, pos, alpha, rad, zc_prev, out, int(use_omp))
    return out


def vring(R: float, core: float, nlayers: int, nazi: int, sig: float):
    """__vring(R, core, nLayers, nAzi, sig) (vring.cpp:50-100) as
    (pos, alpha, radius, volume); alpha is zero until resolved."""
    m = (2 * nlayers + 1) ** 2
    n = nazi * m
    pos, alpha = np.zeros((n, 3)), np.zeros((n, 3))
    radius, volume = np.zeros(n), np.zeros(n)
    lib().or_vring(R, core, nlayers, nazi, sig, pos, alpha, radius, volume)
    return pos, alpha, radius, volume


def vring_ideal(R: float, core: float, nlayers: int) -> np.ndarray:
    out = np.zeros((2 * nlayers + 1) ** 2)
    lib().or_vring_ideal(R, core, nlayers, out)
    return out


def vring_influence(nlayers: int, nazi: int, pos, radius) -> np.ndarray:
    m = (2 * nlayers + 1) ** 2
    M = np.zeros((m, m))
    lib().or_vring_influence(nlayers, nazi, np.ascontiguousarray(pos, dtype=np.float64),
                             np.ascontiguousarray(radius, dtype=np.float64), M)
    return M


def lu_solve(A, b) -> np.ndarray:
    A = np.array(A, dtype=np.float64, order="C")
    x = np.zeros(len(b))
    if lib().or_lu_solve(len(b), A, np.ascontiguousarray(b, dtype=np.float64), x) != 0:
        raise ValueError("singular matrix")
    return x


def vring_resolve(R: float, core: float, nlayers: int, nazi: int, sig: float):
    """__resolve::rebuild on one ring (resolve.cpp:12-63): the strengths of
    the (2L+1)^2 cells and the ring's particles carrying them."""
    pos, alpha, radius, volume = vring(R, core, nlayers, nazi, sig)
    M = vring_influence(nlayers, nazi, pos, radius)
    s = lu_solve(M, vring_ideal(R, core, nlayers))
    m = len(s)
    psi = np.repeat(np.arange(nazi) * (2.0 * np.pi / nazi), m)
    strength = np.tile(s, nazi)
    alpha = np.stack([-np.sin(psi) * strength, np.cos(psi) * strength, np.zeros_like(psi)], axis=1)
    return s, M, (pos, alpha, radius, volume)


class WakeParticles:
    """The particle arrays of one wake as __wake holds them: maxn rows, the
    first `n` in use (positions, strengths: rows of 3)."""

    def __init__(self, pos, alpha, radius, volume, birth=None, active=None, maxn=None):
        n = len(pos)
        maxn = n if maxn is None else maxn
        self.maxn = maxn
        self.n = n
        self.npidx = n
        self.pos = np.zeros((maxn, 3))
        self.alpha = np.zeros((maxn, 3))
        self.radius = np.zeros(maxn)
        self.volume = np.zeros(maxn)
        self.birth = np.zeros(maxn)
        self.active = np.zeros(maxn)
        self.pos[:n], self.alpha[:n] = pos, alpha
        self.radius[:n], self.volume[:n] = radius, volume
        if birth is not None:
            self.birth[:n] = birth
        if active is not None:
            self.active[:n] = active

    def copy(self):
        w = WakeParticles(self.pos[:self.n], self.alpha[:self.n], self.radius[:self.n], self.volume[:self.n],
                          self.birth[:self.n], self.active[:self.n], self.maxn)
        w.npidx = self.npidx
        return w


def wake_split(w: WakeParticles, stepnum: float) -> int:
    """__wake::split (wake.cpp:61-110) in place; the number split, or -1
    when nothing was done (npidx != n, or no room in maxn rows)."""
    n, k = C.c_int64(w.n), C.c_int64(w.npidx)
    r = lib().or_wake_split(float(stepnum), w.maxn, C.byref(n), C.byref(k), w.pos, w.alpha, w.radius,
                            w.volume, w.birth, w.active)
    w.n, w.npidx = n.value, k.value
    return int(r)


def wake_merge(w: WakeParticles, stepnum: float) -> int:
    """__wake::merge (wake.cpp:112-159) in place; the number of merges."""
    return int(lib().or_wake_merge(float(stepnum), w.n, w.pos, w.alpha, w.radius, w.volume, w.active))
