"""CPU oracle of the learned far field (test infrastructure only).

This module is imported by tests/ only.  Nothing here is used by the
library: it restates, in FP64 numpy, what the HIP tree path computes, so
the GPU results can be checked.  The reference has no far-field method (its
solver is the direct sum, src/interaction/interaction.cpp:196-210 with
INTERACT, src/interaction/interaction_utils.h:456-501); the expansion is
pinned here by mathematics instead (tests/test_oracle_farfield.py):

  * md_eval (monopole + dipole of the singular kernel 1/(4 pi r^3)):
    exact for one particle at the centre; its gradient equals the finite
    difference of its velocity; its error against the exact cluster sum
    falls as eps^2 when the cluster shrinks by eps (a dropped or wrong
    dipole term leaves an O(eps) error); the transposed contraction
    alpha^T G equals the reference's pair stretching summed over the cluster
    for a one-particle cluster.
  * quad_eval (the quadrupole term, written with the explicit second and
    third derivative tensors of r/|r|^3): md_eval + quad_eval against the
    exact cluster sum falls as eps^3 (a dropped or wrong quadrupole leaves
    eps^2); its gradient equals the finite difference of its velocity; for
    one particle off the centre it equals half the second directional
    difference of that particle's singular field.
  * morton_perm: the codes of hand-placed points.
  * leaf_first: hand-made codes (a dense cell cut into runs of 256, separate
    cells never sharing a leaf).
  * node_moments: computed from the particles directly (not by the
    parallel-axis shift the GPU uses), so agreement checks the shift.
  * mlp_*: forward and backward of the network, the gradient checked
    against central finite differences.

parity: md_eval, quad_eval, exact_node, morton_perm, node_moments, mlp_forward and
mlp_grad are pinned as above; build_lists is checked through properties
(every source leaf covered exactly once per target leaf; far nodes satisfy
the opening test), not against a reference list.
"""
from __future__ import annotations

import math

import numpy as np

FOUR_PI = 4.0 * math.pi
LEAF = 256
FAN = 8
N_IN, N_HID, N_OUT = 16, 128, 12
OFF = {}
_o = 0
for _name, _size in (("W1", N_IN * N_HID), ("b1", N_HID), ("W2", N_HID * N_HID), ("b2", N_HID),
                     ("W3", N_HID * N_HID), ("b3", N_HID), ("W4", N_HID * N_OUT), ("b4", N_OUT)):
    OFF[_name] = (_o, _size)
    _o += _size
N_PARAMS = _o  # 36,748


# ---------------------------------------------------------------------------
# Tree
# ---------------------------------------------------------------------------
def _spread10(v: np.ndarray) -> np.ndarray:
    v = v.astype(np.uint64) & 0x3FF
    v = (v | (v << 16)) & 0x30000FF
    v = (v | (v << 8)) & 0x300F00F
    v = (v | (v << 4)) & 0x30C30C3
    v = (v | (v << 2)) & 0x9249249
    return v


def morton_codes(pos: np.ndarray) -> np.ndarray:
    """30-bit Morton codes on a 1024^3 grid over the bounding cube."""
    lo, hi = pos.min(axis=0), pos.max(axis=0)
    span = max(float(np.max(hi - lo)), 1e-300)
    s = 1023.0 / (span * 1.0000001)
    q = np.floor((pos - lo) * s).astype(np.int64)
    return _spread10(q[:, 0]) | (_spread10(q[:, 1]) << 1) | (_spread10(q[:, 2]) << 2)


def morton_perm(pos: np.ndarray) -> np.ndarray:
    """Sorted position -> particle index (stable in the particle index)."""
    return np.argsort(morton_codes(pos), kind="stable")


CELL_TARGET = 512


def prefix_levels(codes: np.ndarray) -> np.ndarray:
    """Length in 3-bit groups of the common Morton prefix of sorted
    neighbours i-1, i (i >= 1); 10 for identical codes."""
    c = codes.astype(np.uint64)
    x = c[1:] ^ c[:-1]
    lev = np.full(len(x), 10, dtype=np.int64)
    nz = x != 0
    # bit length of x over the 30-bit code: common prefix = 30 - bitlength
    bl = np.floor(np.log2(x[nz].astype(np.float64))).astype(np.int64) + 1
    lev[nz] = (30 - bl) // 3
    return lev


def leaf_first(codes: np.ndarray) -> np.ndarray:
    """First sorted index of every leaf, then n.  Leaves are the non-empty
    cells of the first octree level whose mean occupancy is at most
    CELL_TARGET, each cut into runs of at most LEAF particles."""
    n = len(codes)
    lev = prefix_levels(codes)
    hist = np.bincount(lev, minlength=11)
    level, cells = 10, 1
    for D in range(1, 11):
        cells += hist[D - 1]
        if n <= CELL_TARGET * cells:
            level = D
            break
    starts = np.concatenate([[0], np.nonzero(lev < level)[0] + 1])
    ends = np.concatenate([starts[1:], [n]])
    firsts = [np.arange(a, b, LEAF) for a, b in zip(starts, ends)]
    return np.concatenate(firsts + [np.array([n])]).astype(np.int64)


def tree_levels(nleaf: int):
    counts = [nleaf]
    while counts[-1] > 1:
        counts.append((counts[-1] + FAN - 1) // FAN)
    offs = np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(int)
    return counts, offs


def node_ranges(lf: np.ndarray):
    """[first, last) of every node from the leaves' firsts (lf, ending in n):
    leaves first, then each level of 8 consecutive nodes."""
    nleaf = len(lf) - 1
    counts, offs = tree_levels(nleaf)
    first, last, level, child = list(lf[:-1]), list(lf[1:]), [0] * nleaf, [-1] * nleaf
    for L in range(1, len(counts)):
        for j in range(counts[L]):
            c0 = offs[L - 1] + FAN * j
            c1 = offs[L - 1] + min(FAN * (j + 1), counts[L - 1])
            first.append(first[c0])
            last.append(last[c1 - 1])
            level.append(L)
            child.append(c0)
    return np.array(first), np.array(last), np.array(level), np.array(child)


def node_moments(pos: np.ndarray, alpha: np.ndarray, sigma: np.ndarray, lf: np.ndarray):
    """Moments of every node straight from its particles (sorted arrays;
    lf = leaf_first of the sorted codes).

    Returns dict of arrays: centre (k,3), r, s0, s1, s2, s3, sg, M (k,3),
    D (k,3,3) with D[i, j] = sum a_i delta_j, Q (k,3,3,3) with
    Q[p, j, l] = sum a_p delta_j delta_l, plus first/last/level/child
    (s_n = sum |a| |delta|^n).
    """
    first, last, level, child = node_ranges(lf)
    k = len(first)
    out = {"centre": np.zeros((k, 3)), "r": np.zeros(k), "s0": np.zeros(k), "s1": np.zeros(k),
           "s2": np.zeros(k), "s3": np.zeros(k), "sg": np.zeros(k), "M": np.zeros((k, 3)),
           "D": np.zeros((k, 3, 3)), "Q": np.zeros((k, 3, 3, 3)),
           "first": first, "last": last, "level": level, "child": child}
    am = np.linalg.norm(alpha, axis=1)
    for i in range(k):
        a, b = first[i], last[i]
        c = pos[a:b].mean(axis=0)
        d = pos[a:b] - c
        out["centre"][i] = c
        out["r"][i] = np.sqrt((d * d).sum(axis=1)).max()
        out["s0"][i] = am[a:b].sum()
        dn = np.sqrt((d * d).sum(axis=1))
        out["s1"][i] = (am[a:b] * dn).sum()
        out["s2"][i] = (am[a:b] * dn * dn).sum()
        out["s3"][i] = (am[a:b] * dn ** 3).sum()
        out["sg"][i] = sigma[a:b].mean()
        out["M"][i] = alpha[a:b].sum(axis=0)
        out["D"][i] = alpha[a:b].T @ d
        out["Q"][i] = np.einsum("sp,sj,sl->pjl", alpha[a:b], d, d)
    return out


def list_partition_errors(nodes_first, nodes_last, nleaf, near_cnt, near, far_cnt, far) -> int:
    """Number of target leaves whose near leaves and far nodes do not
    cover every particle exactly once."""
    bad = 0
    n = int(nodes_last[nleaf - 1])
    for T in range(nleaf):
        cover = np.zeros(n, dtype=np.int32)
        for id_ in near[T, :near_cnt[T]]:
            cover[nodes_first[id_]:nodes_last[id_]] += 1
        for id_ in far[T, :far_cnt[T]]:
            cover[nodes_first[id_]:nodes_last[id_]] += 1
        if not np.all(cover == 1):
            bad += 1
    return bad


# ---------------------------------------------------------------------------
# Far-field expansion and exact node sums (4 pi factored: u' = 4 pi u)
# ---------------------------------------------------------------------------
def md_eval(r: np.ndarray, M: np.ndarray, D: np.ndarray):
    """Monopole + dipole of the singular kernel at offsets r (k,3) from the
    centre: returns u' (k,3) and G' (k,3,3), G'[k][m] = du'_k/dx_m."""
    r = np.atleast_2d(r)
    r2 = (r * r).sum(axis=1)
    ir = 1.0 / np.sqrt(r2)
    ir3, ir5, ir7 = ir ** 3, ir ** 5, ir ** 7
    mr = np.cross(M, r)
    w = np.array([D[1, 2] - D[2, 1], D[2, 0] - D[0, 2], D[0, 1] - D[1, 0]])
    p = r @ D.T  # p_i = D_ij r_j
    pr = np.cross(p, r)
    u = mr * ir3[:, None] - w[None, :] * ir3[:, None] + 3.0 * pr * ir5[:, None]
    k = len(r)
    G = np.zeros((k, 3, 3))
    cm = np.array([[0, -M[2], M[1]], [M[2], 0, -M[0]], [-M[1], M[0], 0]])
    G += cm[None] * ir3[:, None, None]
    # 3 [p]x / r^5
    px = np.zeros((k, 3, 3))
    px[:, 0, 1], px[:, 0, 2] = -p[:, 2], p[:, 1]
    px[:, 1, 0], px[:, 1, 2] = p[:, 2], -p[:, 0]
    px[:, 2, 0], px[:, 2, 1] = -p[:, 1], p[:, 0]
    G += 3.0 * px * ir5[:, None, None]
    # 3 (D[:, m] x r) / r^5
    for m in range(3):
        G[:, :, m] += 3.0 * np.cross(np.broadcast_to(D[:, m], r.shape), r) * ir5[:, None]
    v = 3.0 * (w[None, :] - mr) * ir5[:, None] - 15.0 * pr * ir7[:, None]
    G += v[:, :, None] * r[:, None, :]
    return u, G


def _eps():
    e = np.zeros((3, 3, 3))
    e[0, 1, 2] = e[1, 2, 0] = e[2, 0, 1] = 1.0
    e[0, 2, 1] = e[2, 1, 0] = e[1, 0, 2] = -1.0
    return e


QUAD_SLOTS = ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))


def quad_from_slots(q18: np.ndarray) -> np.ndarray:
    """Q[p, j, l] (3,3,3) from the 18 exported values: for each component
    p, the pairs (j, l) in the order xx yy zz xy xz yz."""
    Q = np.zeros((3, 3, 3))
    for p in range(3):
        for m, (j, l) in enumerate(QUAD_SLOTS):
            Q[p, j, l] = Q[p, l, j] = q18[6 * p + m]
    return Q


def quad_eval(r: np.ndarray, Q: np.ndarray):
    """Quadrupole term of the singular kernel at offsets r (k,3) from the
    centre, Q[p, j, l] = sum a_p delta_j delta_l: u'_k = (1/2) eps_kpi
    Q[p,j,l] d_j d_l K_i(r), K = r/|r|^3, and G'_km = du'_k/dx_m with the
    third derivatives of K.  Returns u' (k,3), G' (k,3,3)."""
    r = np.atleast_2d(r)
    eye = np.eye(3)
    eps = _eps()
    us, Gs = [], []
    for x in r:
        r2 = x @ x
        i5, i7, i9 = r2 ** -2.5, r2 ** -3.5, r2 ** -4.5
        # d_j d_l K_i
        T2 = (-3.0 * (np.einsum("ij,l->ijl", eye, x) + np.einsum("il,j->ijl", eye, x)
                      + np.einsum("jl,i->ijl", eye, x)) * i5
              + 15.0 * np.einsum("i,j,l->ijl", x, x, x) * i7)
        # d_m d_j d_l K_i
        T3 = (-3.0 * (np.einsum("ij,lm->ijlm", eye, eye) + np.einsum("il,jm->ijlm", eye, eye)
                      + np.einsum("jl,im->ijlm", eye, eye)) * i5
              + 15.0 * (np.einsum("ij,l,m->ijlm", eye, x, x) + np.einsum("il,j,m->ijlm", eye, x, x)
                        + np.einsum("jl,i,m->ijlm", eye, x, x) + np.einsum("im,j,l->ijlm", eye, x, x)
                        + np.einsum("jm,i,l->ijlm", eye, x, x) + np.einsum("lm,i,j->ijlm", eye, x, x)) * i7
              - 105.0 * np.einsum("i,j,l,m->ijlm", x, x, x, x) * i9)
        us.append(0.5 * np.einsum("kpi,pjl,ijl->k", eps, Q, T2))
        Gs.append(0.5 * np.einsum("kpi,pjl,ijlm->km", eps, Q, T3))
    return np.array(us), np.array(Gs)


def exact_node(x: np.ndarray, sig_t: float, pos: np.ndarray, alpha: np.ndarray, sigma: np.ndarray):
    """u', G' at target x of the particles (pos, alpha, sigma) with the
    reference's Gaussian kernel (gaussian.h:47-60): q = (phi/rho_bar - Z) /
    rho_bar^2, F = (Z - 3 q) / rho^2."""
    d = x[None, :] - pos
    r2 = (d * d).sum(axis=1)
    sig2 = 0.5 * (sigma ** 2 + sig_t ** 2)
    sig = np.sqrt(sig2)
    rb = np.sqrt(r2) / sig
    from scipy.special import erf
    Z = (2.0 * math.pi) ** -1.5 * np.exp(-0.5 * rb * rb) / (sig2 * sig)
    phi = erf(rb / math.sqrt(2.0)) / (FOUR_PI * sig2 * sig)
    q = (phi / rb - Z) / (rb * rb)
    F = (Z - 3.0 * q) / r2
    ad = np.cross(alpha, d)
    u = (q[:, None] * ad).sum(axis=0)
    G = np.zeros((3, 3))
    for k in range(3):
        for m in range(3):
            G[k, m] = (F * d[:, m] * ad[:, k]).sum()
    # q eps_kim a_i
    aq = (q[:, None] * alpha).sum(axis=0)
    G += np.array([[0, -aq[2], aq[1]], [aq[2], 0, -aq[0]], [-aq[1], aq[0], 0]])
    return FOUR_PI * u, FOUR_PI * G


def frame(r: np.ndarray, M: np.ndarray, s0: float) -> np.ndarray:
    """Rows e1, e2, e3 of the network's frame: e1 = M/|M| (r/|r| when
    |M| < 1e-3 s0), e2 = the part of r normal to e1 (any normal if r || e1),
    e3 = e1 x e2."""
    r = np.asarray(r, dtype=np.float64)
    if M @ M > 1e-6 * s0 * s0:
        e1 = M / np.linalg.norm(M)
    else:
        e1 = r / np.linalg.norm(r)
    v = r - (r @ e1) * e1
    if not (v @ v > 1e-12 * (r @ r)):
        a = np.array([1.0, 0.0, 0.0]) if abs(e1[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
        v = a - (a @ e1) * e1
    e2 = v / np.linalg.norm(v)
    return np.array([e1, e2, np.cross(e1, e2)])


def features(r: np.ndarray, ell: float, s0: float, M: np.ndarray, D: np.ndarray, sg: float):
    """Network inputs of offsets r (k,3) from a node, in the frame of each
    offset, with the frames and output scales: (X, su, sG, R); the scales
    are those of the first term monopole + dipole leave out (the
    quadrupole), su = s0 l^2 / r^4, sG = su / r."""
    r = np.atleast_2d(r)
    d2 = (r * r).sum(axis=1)
    X = np.zeros((len(r), N_IN))
    Rs = np.zeros((len(r), 3, 3))
    for i in range(len(r)):
        R = frame(r[i], M, s0)
        Rs[i] = R
        X[i, 0:3] = R @ r[i] * (ell / d2[i])
        X[i, 3:6] = R @ M / s0
        X[i, 6:15] = (R @ D @ R.T).reshape(-1) / (s0 * ell)
        X[i, 15] = sg / ell
    su = s0 * ell * ell / (d2 * d2)
    sG = su / np.sqrt(d2)
    return X, su, sG, Rs


# ---------------------------------------------------------------------------
# Network
# ---------------------------------------------------------------------------
def unpack(params: np.ndarray):
    p = np.asarray(params, dtype=np.float64)
    g = {}
    for name, (o, size) in OFF.items():
        g[name] = p[o:o + size]
    g["W1"] = g["W1"].reshape(N_IN, N_HID)
    g["W2"] = g["W2"].reshape(N_HID, N_HID)
    g["W3"] = g["W3"].reshape(N_HID, N_HID)
    g["W4"] = g["W4"].reshape(N_HID, N_OUT)
    return g


C_GELU = math.sqrt(2.0 / math.pi)


def gelu(x):
    """The tanh form of GELU, 0.5 x (1 + tanh z), z = sqrt(2/pi)(x + 0.044715
    x^3), written as x sigmoid(2 z) (the same function without the
    cancellation of 1 + tanh z for negative x)."""
    z = C_GELU * (x + 0.044715 * x ** 3)
    return x / (1.0 + np.exp(-2.0 * z))


def gelu_d(x):
    z = C_GELU * (x + 0.044715 * x ** 3)
    s = 1.0 / (1.0 + np.exp(-2.0 * z))
    return s + x * s * (1.0 - s) * 2.0 * C_GELU * (1.0 + 3.0 * 0.044715 * x * x)


def bf16(x):
    """Round to BF16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def mlp_forward(params: np.ndarray, X: np.ndarray, bf16_emulate: bool = False) -> np.ndarray:
    """FP64 forward; with bf16_emulate, inputs, weights and the activations
    fed to each layer are rounded to BF16 as the inference kernel does."""
    g = unpack(params)
    rnd = bf16 if bf16_emulate else (lambda v: np.asarray(v, dtype=np.float64))
    h = rnd(X)
    h = rnd(gelu(h @ rnd(g["W1"]) + g["b1"]))
    h = rnd(gelu(h @ rnd(g["W2"]) + g["b2"]))
    h = rnd(gelu(h @ rnd(g["W3"]) + g["b3"]))
    return h @ rnd(g["W4"]) + g["b4"]


def mlp_loss_grad(params: np.ndarray, X: np.ndarray, Y: np.ndarray):
    """L = (1/n) sum (f(x) - y)^2 and dL/dparams (FP64)."""
    g = unpack(params)
    n = len(X)
    a1 = X @ g["W1"] + g["b1"]
    h1 = gelu(a1)
    a2 = h1 @ g["W2"] + g["b2"]
    h2 = gelu(a2)
    a3 = h2 @ g["W3"] + g["b3"]
    h3 = gelu(a3)
    out = h3 @ g["W4"] + g["b4"]
    err = out - Y
    loss = float((err * err).sum() / n)
    d4 = 2.0 * err / n
    grads = {"W4": h3.T @ d4, "b4": d4.sum(axis=0)}
    d3 = (d4 @ g["W4"].T) * gelu_d(a3)
    grads["W3"], grads["b3"] = h2.T @ d3, d3.sum(axis=0)
    d2 = (d3 @ g["W3"].T) * gelu_d(a2)
    grads["W2"], grads["b2"] = h1.T @ d2, d2.sum(axis=0)
    d1 = (d2 @ g["W2"].T) * gelu_d(a1)
    grads["W1"], grads["b1"] = X.T @ d1, d1.sum(axis=0)
    flat = np.zeros(N_PARAMS)
    for name, (o, size) in OFF.items():
        flat[o:o + size] = np.asarray(grads[name]).reshape(-1)
    return loss, flat
