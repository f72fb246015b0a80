"""The tree path (PW_LEARNED) with the high-order algebraic kernel
(highorder.h:20-106, selected everywhere by interaction_utils.h:22-26): the
near field sums the kernel's own pairs, and a node is far only when the
kernel's deviation from the singular law the far field expands, at most
(15/8)(sigma/r)^4 of it, fits in the node's error budget.  Checked against
the oracle on the GPU's own lists and against the FP64 direct sum of the
same kernel."""
import math

import numpy as np
import pytest

from conftest import normwise
from oracle import farfield as FF
from pawan_amd import capi, workloads as W

pytestmark = pytest.mark.gpu

HO = capi.PW_KERNEL_HIGHORDER


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def helix():
    return W.helix_wake(32768, seed=3)


@pytest.fixture(scope="module")
def direct(helix):
    """FP64 direct sums of the helix, high-order kernel and Gaussian."""
    out = {}
    with capi.Solver(helix) as g:
        for k in (capi.PW_KERNEL_GAUSSIAN, HO):
            g.set_kernel(k)
            out[k] = g.solve("fp64")
    return out


def _build(helix, kernel, tol):
    with capi.Solver(helix) as g:
        g.set_kernel(kernel)
        g.ff_set_params(0.5, tol, use_mlp=False)
        g.ff_set_conserve(False)
        g.ff_build()
        out = {k: g.ff_export(w) for k, w in (("perm", capi.FF_PERM), ("nodes", capi.FF_NODES),
                                              ("near_cnt", capi.FF_NEAR_CNT), ("near", capi.FF_NEAR),
                                              ("far_cnt", capi.FF_FAR_CNT), ("far", capi.FF_FAR),
                                              ("quad", capi.FF_QUAD))}
        out["info"] = g.ff_info()
        out["coef"] = g.ff_err_coef()[0]
        out["u"], out["w"] = g.solve("learned")
    return out


def _urms(u):
    return float(np.sqrt((u * u).sum(axis=1).mean()))


def test_far_nodes_keep_the_kernel_deviation_within_the_tolerance(helix, direct):
    tol = 1e-4 * _urms(direct[HO][0])
    b = _build(helix, HO, tol)
    nd, nl = b["nodes"], b["info"]["leaves"]
    first, last = nd[:, 19].astype(np.int64), nd[:, 20].astype(np.int64)
    assert b["info"]["overflow"] == 0
    assert FF.list_partition_errors(first, last, nl, b["near_cnt"], b["near"], b["far_cnt"], b["far"]) == 0
    s4 = float(np.max(helix.radius) ** 4)
    nfar, worst = 0, 0.0
    for T in range(nl):
        for q in b["far"][T, :b["far_cnt"][T]]:
            d = float(np.linalg.norm(nd[q, 0:3] - nd[T, 0:3]))
            gap, dm = d - nd[q, 3] - nd[T, 3], d - nd[T, 3]
            e = b["coef"] * b["quad"][q, 19] / dm ** 5 + 1.875 * s4 * nd[q, 4] / (4 * math.pi * gap ** 6)
            worst = max(worst, e / tol)
            nfar += 1
    _say(f"high-order lists at tol {tol:.3e}: {nfar} far nodes, worst budget use {worst:.4f}, "
         f"near pairs {b['info']['near_pairs']:.3e}")
    assert nfar > 1000
    assert worst <= 1.0 + 1e-4  # FP32 evaluation of the same test


def test_near_field_matches_the_oracle_on_gpu_lists(oracle, helix, direct):
    """Near leaves by the oracle's high-order pair sums, far nodes by the
    expansion, at sampled targets (the FP32C bars of the Gaussian path)."""
    b = _build(helix, HO, 1e-4 * _urms(direct[HO][0]))
    perm, nd = b["perm"], b["nodes"]
    sp, sa = helix.pos[perm], helix.alpha[perm]
    sr, sv = helix.radius[perm], helix.volume[perm]
    nl = b["info"]["leaves"]
    rng = np.random.default_rng(0)
    targets = np.sort(rng.choice(helix.n, size=40, replace=False))
    ru, rw = [], []
    with oracle.smoothing(oracle.HIGHORDER):
        for t in targets:
            T = int(np.searchsorted(nd[:nl, 19], t, side="right")) - 1
            src = []
            for q in b["near"][T, :b["near_cnt"][T]]:
                src.extend(range(int(nd[q, 19]), int(nd[q, 20])))
            idx = np.array([t] + [s for s in src if s != t])
            sub = W.ParticleSystem(off=np.array([0, len(idx)]), pos=sp[idx], alpha=sa[idx], radius=sr[idx],
                                   volume=sv[idx], nu=helix.nu, coupled=True, name="near")
            u, w = oracle.rates_targets(sub, [0])
            u, w = u[0].copy(), w[0].copy()
            for q in b["far"][T, :b["far_cnt"][T]]:
                r = (sp[t] - nd[q, 0:3])[None]
                um, Gm = FF.md_eval(r, nd[q, 7:10], nd[q, 10:19].reshape(3, 3))
                uq, Gq = FF.quad_eval(r, FF.quad_from_slots(b["quad"][q, :18]))
                u += (um[0] + uq[0]) / FF.FOUR_PI
                w += sa[t] @ (Gm[0] + Gq[0]) / FF.FOUR_PI
            ru.append(u)
            rw.append(w)
    gu, gw = b["u"][perm[targets]], b["w"][perm[targets]]
    eu, ew = normwise(gu, np.array(ru)), normwise(gw, np.array(rw))
    _say(f"high-order tree path vs oracle on its lists: e_u={eu:.2e} e_w={ew:.2e}")
    assert eu < 1e-5 and ew < 1e-5


def test_error_against_the_direct_sum_falls_with_the_tolerance(helix, direct):
    """e = max_i |du_i| / max_i |u_i| against the FP64 direct sum of the
    same kernel falls with the tolerance; the high-order lists are never
    shorter in near pairs than the Gaussian's at the same tolerance (the
    deviation term only adds to a node's error budget)."""
    ru, rw = direct[HO]
    urms = _urms(ru)
    errs, near = [], []
    for k in (1e-3, 1e-4, 1e-5):
        b = _build(helix, HO, k * urms)
        bg = _build(helix, capi.PW_KERNEL_GAUSSIAN, k * urms)
        e = float(np.linalg.norm(b["u"] - ru, axis=1).max() / np.linalg.norm(ru, axis=1).max())
        ew = normwise(b["w"], rw)
        errs.append(e)
        near.append((b["info"]["near_pairs"], bg["info"]["near_pairs"]))
        _say(f"tol {k:.0e} urms: e_max={e:.3e} e_w={ew:.3e} near pairs {near[-1][0]:.3e} "
             f"(Gaussian {near[-1][1]:.3e})")
        assert np.all(np.isfinite(b["u"])) and np.all(np.isfinite(b["w"]))
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] < 1e-4
    assert all(h >= g for h, g in near)


def test_without_a_tolerance_far_nodes_lie_beyond_the_deviation_radius(helix):
    """tol = 0: the opening angle alone for the Gaussian; the high-order
    kernel also needs gap^2 > 5600 sigma_max^2 (deviation below 2^-24)."""
    b = _build(helix, HO, 0.0)
    nd, nl = b["nodes"], b["info"]["leaves"]
    s2 = float(np.max(helix.radius) ** 2)
    for T in range(nl):
        for q in b["far"][T, :b["far_cnt"][T]]:
            gap = float(np.linalg.norm(nd[q, 0:3] - nd[T, 0:3])) - nd[q, 3] - nd[T, 3]
            assert gap * gap > 5600 * s2 * (1 - 1e-5)
    first, last = nd[:, 19].astype(np.int64), nd[:, 20].astype(np.int64)
    assert FF.list_partition_errors(first, last, nl, b["near_cnt"], b["near"], b["far_cnt"], b["far"]) == 0
