"""Seeded synthetic particle systems for the five benchmark workloads.

Every generator returns a :class:`ParticleSystem`: wakes concatenated in one
set of arrays (positions and vorticity strengths ``N x 3`` row-major, like the
rows of the reference's ``__wake::_position``/``_vorticity`` matrices,
src/wake/wake.h:49-52; smoothing radii and volumes ``N``), with ``off`` the
wake offsets.

Conventions reused from the reference's initial conditions:
  * tangential strength ``Gamma * R * dpsi`` along ``(-sin psi, cos psi, 0)``
    and volume ``4 pi sigma^3 / 3`` for a thin ring (src/wake/ring.cpp:22-39);
  * cross-section positions ``((R + r cos th) cos psi, (R + r cos th) sin psi,
    r sin th)`` (src/wake/vring.cpp:100-102).

Random draws use splitmix64 (Steele, Lea & Flood, "Fast splittable
pseudorandom number generators", OOPSLA 2014) in counter mode, so any draw
is a pure function of (seed, stream, index) and is identical on every
machine.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)
_GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(seed: int, stream: int, idx: np.ndarray) -> np.ndarray:
    """splitmix64 output for counters ``idx`` of (seed, stream)."""
    with np.errstate(over="ignore"):
        base = np.uint64((seed * 0x9E3779B97F4A7C15 + stream * 0xD1B54A32D192ED03) & 0xFFFFFFFFFFFFFFFF)
        z = base + (np.asarray(idx, dtype=np.uint64) + np.uint64(1)) * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, stream: int, n: int, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
    """``n`` doubles in [lo, hi) from the top 53 bits of splitmix64."""
    z = splitmix64(seed, stream, np.arange(n, dtype=np.uint64))
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return lo + (hi - lo) * u


@dataclasses.dataclass
class ParticleSystem:
    """Particles of one or more wakes, concatenated.

    ``coupled`` is True when all wakes interact (the reference's
    ``__interaction`` over several wakes, src/interaction/interaction.cpp:
    196-210) and False when each wake is an independent system (the batched
    workload: one ``__interaction`` per wake).
    """

    off: np.ndarray  # int64 [nw + 1]
    pos: np.ndarray  # float64 [N, 3]
    alpha: np.ndarray  # float64 [N, 3]
    radius: np.ndarray  # float64 [N]
    volume: np.ndarray  # float64 [N]
    nu: float = 0.0
    coupled: bool = True
    name: str = ""

    @property
    def n(self) -> int:
        return int(self.off[-1])

    @property
    def nw(self) -> int:
        return len(self.off) - 1

    def copy(self) -> "ParticleSystem":
        return dataclasses.replace(
            self,
            off=self.off.copy(),
            pos=self.pos.copy(),
            alpha=self.alpha.copy(),
            radius=self.radius.copy(),
            volume=self.volume.copy(),
        )

    def wake(self, w: int) -> "ParticleSystem":
        a, b = int(self.off[w]), int(self.off[w + 1])
        return ParticleSystem(
            off=np.array([0, b - a], dtype=np.int64),
            pos=self.pos[a:b].copy(),
            alpha=self.alpha[a:b].copy(),
            radius=self.radius[a:b].copy(),
            volume=self.volume[a:b].copy(),
            nu=self.nu,
            coupled=True,
            name=f"{self.name}[{w}]",
        )

    def ordered_pairs(self) -> int:
        """Ordered (target, source) interactions per right-hand side."""
        sizes = np.diff(self.off).astype(np.int64)
        if self.coupled:
            return self.n * (self.n - 1)
        return int(np.sum(sizes * (sizes - 1)))


def concat(parts: list[ParticleSystem], coupled: bool, nu: float, name: str) -> ParticleSystem:
    sizes = [p.n for p in parts]
    off = np.zeros(len(parts) + 1, dtype=np.int64)
    off[1:] = np.cumsum(sizes)
    return ParticleSystem(
        off=off,
        pos=np.ascontiguousarray(np.concatenate([p.pos for p in parts])),
        alpha=np.ascontiguousarray(np.concatenate([p.alpha for p in parts])),
        radius=np.concatenate([p.radius for p in parts]),
        volume=np.concatenate([p.volume for p in parts]),
        nu=nu,
        coupled=coupled,
        name=name,
    )


def cored_ring(R: float, gamma: float, sigma: float, core: float, nst: int,
               ncs_ring: int, z0: float = 0.0) -> ParticleSystem:
    """Ring of ``nst`` stations, each a centre particle plus ``ncs_ring``
    particles at radius ``core`` in the cross-section (angles 2 pi k/ncs_ring).

    The station circulation ``gamma * R * dpsi`` (ring.cpp:25) is split by
    Gaussian weights exp(-r^2 / 2 core^2); volumes are pi core^2 R dpsi / ncs.
    """
    ncs = 1 + ncs_ring
    dpsi = 2.0 * math.pi / nst
    psi = np.arange(nst) * dpsi
    r = np.concatenate([[0.0], np.full(ncs_ring, core)])
    th = np.concatenate([[0.0], 2.0 * math.pi * np.arange(ncs_ring) / ncs_ring])
    wgt = np.exp(-0.5 * (r / core) ** 2) if core > 0 else np.ones(ncs)
    wgt = wgt / wgt.sum()
    P, C = np.meshgrid(psi, np.arange(ncs), indexing="ij")
    rr, tt = r[C], th[C]
    cp, sp = np.cos(P), np.sin(P)
    pos = np.stack([(R + rr * np.cos(tt)) * cp, (R + rr * np.cos(tt)) * sp,
                    z0 + rr * np.sin(tt)], axis=-1).reshape(-1, 3)
    mag = gamma * R * dpsi * wgt[C]
    alpha = np.stack([-mag * sp, mag * cp, np.zeros_like(mag)], axis=-1).reshape(-1, 3)
    n = nst * ncs
    vol = np.full(n, math.pi * core * core * R * dpsi / ncs)
    return ParticleSystem(off=np.array([0, n], dtype=np.int64), pos=np.ascontiguousarray(pos),
                          alpha=np.ascontiguousarray(alpha), radius=np.full(n, sigma), volume=vol)


def thin_ring(R: float, gamma: float, sigma: float, n: int) -> ParticleSystem:
    """Thin ring exactly as ``__ring(gamma, radius, core, nRadial)``
    (src/wake/ring.cpp:22-39): one particle per station, strength
    gamma*R*dpsi, volume 4 pi core^3 / 3."""
    dpsi = 2.0 * math.pi / n
    psi = np.arange(n) * dpsi
    pos = np.stack([R * np.cos(psi), R * np.sin(psi), np.zeros(n)], axis=-1)
    s = gamma * R * dpsi
    alpha = np.stack([-s * np.sin(psi), s * np.cos(psi), np.zeros(n)], axis=-1)
    return ParticleSystem(off=np.array([0, n], dtype=np.int64), pos=pos, alpha=alpha,
                          radius=np.full(n, sigma), volume=np.full(n, 4.0 * math.pi * sigma ** 3 / 3.0))


# ---------------------------------------------------------------------------
# The five workloads
# ---------------------------------------------------------------------------

def ring_1k() -> ParticleSystem:
    """Single ring, R=1, Gamma=1, sigma=0.1: 200 stations x 5 (centre + 4 at
    r=0.1, theta=k pi/2), N=1000, inviscid.  RK4 dt=0.005 x 100."""
    s = cored_ring(R=1.0, gamma=1.0, sigma=0.1, core=0.1, nst=200, ncs_ring=4)
    s.nu = 0.0
    s.name = "ring_1k"
    return s


def leapfrog_64k(nst: int = 4096) -> ParticleSystem:
    """Two coaxial rings as two coupled wakes: R=1 at z=0 and R=1.2 at z=0.5,
    sigma = core = 0.08 R, Gamma=1, nu=1e-4; ``nst`` stations x 8 (centre + 7
    at r=sigma) each, N = 16 nst (65,536 at the default).  RK4 dt=0.002."""
    parts = []
    for R, z0 in ((1.0, 0.0), (1.2, 0.5)):
        sig = 0.08 * R
        parts.append(cored_ring(R=R, gamma=1.0, sigma=sig, core=sig, nst=nst, ncs_ring=7, z0=z0))
    return concat(parts, coupled=True, nu=1e-4, name=f"leapfrog_{16 * nst}")


def _orthobasis(nrm: np.ndarray):
    nrm = nrm / np.linalg.norm(nrm)
    ref = np.array([1.0, 0.0, 0.0]) if abs(nrm[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    e1 = np.cross(nrm, ref)
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(nrm, e1)
    return e1, e2


def batched_rings(nwakes: int = 256, seed: int = 2024, nmin: int = 512, nmax: int = 16384,
                  nu: float = 1e-4) -> ParticleSystem:
    """``nwakes`` independent thin rings (ring.cpp:22-39 each), N_i =
    round(exp(U[ln nmin, ln nmax])), R ~ U[0.5, 2], sigma/R ~ U[0.05, 0.2],
    Gamma ~ U[0.5, 2], normal uniform on the sphere, centre ~ U[-50, 50]^3."""
    u = lambda stream, lo, hi: uniform(seed, stream, nwakes, lo, hi)
    ns = np.rint(np.exp(u(1, math.log(nmin), math.log(nmax)))).astype(np.int64)
    Rs, sr, gs = u(2, 0.5, 2.0), u(3, 0.05, 0.2), u(4, 0.5, 2.0)
    cz, ph = u(5, -1.0, 1.0), u(6, 0.0, 2.0 * math.pi)
    cen = np.stack([u(7, -50, 50), u(8, -50, 50), u(9, -50, 50)], axis=-1)
    parts = []
    for w in range(nwakes):
        st = math.sqrt(max(0.0, 1.0 - cz[w] ** 2))
        nrm = np.array([st * math.cos(ph[w]), st * math.sin(ph[w]), cz[w]])
        e1, e2 = _orthobasis(nrm)
        ring = thin_ring(float(Rs[w]), float(gs[w]), float(sr[w] * Rs[w]), int(ns[w]))
        # rotate the z=0 ring into the (e1, e2) plane and move it to its centre
        M = np.stack([e1, e2, nrm], axis=-1)
        ring.pos = np.ascontiguousarray(ring.pos @ M.T + cen[w])
        ring.alpha = np.ascontiguousarray(ring.alpha @ M.T)
        parts.append(ring)
    return concat(parts, coupled=False, nu=nu, name=f"batched_{nwakes}")


def helix_wake(n: int = 1 << 20, seed: int = 7, nu: float = 1.56e-5) -> ParticleSystem:
    """Rotor-like wake: 4 contracting tip helices plus a jittered inboard sheet.

    Half the particles lie on 4 tip helices (blade phase b pi/2): psi in
    [0, 8 pi) (4 revolutions), radius r(psi) = 1 - 0.2 psi / (8 pi) (contracts
    to 0.8 over 4 revs), z(psi) = -0.1 psi / (2 pi) (pitch 0.1 per rev),
    circulation 1, strength = Gamma * dx/dpsi * dpsi.  The other half form
    the inboard sheet: per blade 16 helical strips at radius fractions
    f_j = 0.2 + 0.75 (j + 1/2)/16 of r(psi), each of circulation -0.5/16,
    with every particle jittered uniformly within its psi step and its strip
    width.  sigma = 0.02, volume = pi sigma^2 ds.  nu defaults to the
    reference's hard-coded rotor-wake value 1.56e-5 (interaction.cpp:15).
    """
    if n % 128:
        raise ValueError("helix_wake needs n divisible by 128")
    sigma = 0.02
    ntip = n // 8  # per blade
    nstrip = 16
    nsheet = n // 2 // (4 * nstrip)  # per strip
    rev = 4
    psimax = 2.0 * math.pi * rev
    drdpsi, dzdpsi = -0.2 / psimax, -0.1 / (2.0 * math.pi)

    def helix(psi, frac, phase, dpsi, gamma):
        r = (1.0 - 0.2 * psi / psimax) * frac
        th = psi + phase
        c, s = np.cos(th), np.sin(th)
        pos = np.stack([r * c, r * s, dzdpsi * psi], axis=-1)
        dr = drdpsi * frac
        tan = np.stack([dr * c - r * s, dr * s + r * c, np.full_like(psi, dzdpsi)], axis=-1)
        alpha = gamma * tan * dpsi
        ds = np.linalg.norm(tan, axis=-1) * dpsi
        return pos, alpha, ds

    P, A, V = [], [], []
    for b in range(4):
        dpsi = psimax / ntip
        psi = np.arange(ntip) * dpsi
        pos, alpha, ds = helix(psi, 1.0, b * math.pi / 2, dpsi, 1.0)
        P.append(pos), A.append(alpha), V.append(ds)
    width = 0.75 / nstrip
    for b in range(4):
        for j in range(nstrip):
            stream = 100 + 2 * (b * nstrip + j)
            dpsi = psimax / nsheet
            psi = (np.arange(nsheet) + uniform(seed, stream, nsheet)) * dpsi
            frac = 0.2 + width * (j + uniform(seed, stream + 1, nsheet))
            pos, alpha, ds = helix(psi, frac, b * math.pi / 2, dpsi, -0.5 / nstrip)
            P.append(pos), A.append(alpha), V.append(ds)
    pos = np.ascontiguousarray(np.concatenate(P))
    alpha = np.ascontiguousarray(np.concatenate(A))
    ds = np.concatenate(V)
    return ParticleSystem(off=np.array([0, n], dtype=np.int64), pos=pos, alpha=alpha,
                          radius=np.full(n, sigma), volume=math.pi * sigma * sigma * ds,
                          nu=nu, coupled=True, name=f"helix_{n}")


WORKLOADS = {
    "ring_1k": dict(make=ring_1k, dt=0.005, steps=100),
    "leapfrog_64k": dict(make=leapfrog_64k, dt=0.002, steps=200),
    "batched_256": dict(make=batched_rings, dt=0.002, steps=200),
    "helix_1m": dict(make=lambda: helix_wake(1 << 20), dt=1e-3, steps=50),
    "helix_2m": dict(make=lambda: helix_wake(1 << 21), dt=1e-3, steps=200),
}
