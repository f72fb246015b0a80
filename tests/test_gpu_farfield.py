"""The tree path (PW_LEARNED) on the GPU against the far-field oracle:
Morton order, node moments, the interaction lists, the analytic far field
on the GPU's own lists, training samples, the network's inference (BF16
matrix cores) and training gradient, and the approximation error against
the direct sum."""
import numpy as np
import pytest

from conftest import normwise
from oracle import farfield as FF
from pawan_amd import capi, workloads as W

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def helix():
    return W.helix_wake(32768, seed=3)


@pytest.fixture(scope="module")
def built(helix):
    """Tree, lists and exports of the helix at theta 0.5 (no error test)."""
    with capi.Solver(helix) as g:
        g.ff_set_params(0.5, 0.0, use_mlp=False)
        g.ff_set_conserve(False)  # the oracle restates the raw far field
        g.ff_build()
        out = {k: g.ff_export(w) for k, w in (("perm", capi.FF_PERM), ("nodes", capi.FF_NODES),
                                              ("near_cnt", capi.FF_NEAR_CNT), ("near", capi.FF_NEAR),
                                              ("far_cnt", capi.FF_FAR_CNT), ("far", capi.FF_FAR),
                                              ("quad", capi.FF_QUAD))}
        out["info"] = g.ff_info()
        u, w = g.solve("learned")
        out["u"], out["w"] = u, w
    _say(f"tree {out['info']}")
    return out


def test_morton_order_matches_oracle(helix, built):
    np.testing.assert_array_equal(built["perm"], FF.morton_perm(helix.pos))


def test_leaves_match_oracle(helix, built):
    lf = FF.leaf_first(FF.morton_codes(helix.pos)[built["perm"]])
    nd, nl = built["nodes"], built["info"]["leaves"]
    assert nl == len(lf) - 1
    np.testing.assert_array_equal(nd[:nl, 19], lf[:-1])
    np.testing.assert_array_equal(nd[:nl, 20], lf[1:])


def test_node_moments_match_oracle(helix, built):
    perm = built["perm"]
    lf = FF.leaf_first(FF.morton_codes(helix.pos)[perm])
    mo = FF.node_moments(helix.pos[perm], helix.alpha[perm], helix.radius[perm], lf)
    nd = built["nodes"]
    ext = np.ptp(helix.pos, axis=0).max()
    np.testing.assert_array_equal(nd[:, 19], mo["first"])
    np.testing.assert_array_equal(nd[:, 20], mo["last"])
    np.testing.assert_array_equal(nd[:, 21], mo["level"])
    np.testing.assert_allclose(nd[:, 0:3], mo["centre"], rtol=0, atol=1e-12 * ext)
    leaf = mo["level"] == 0
    # radius: rounded-up FP32 bound; exact for leaves, a bound for parents
    assert np.all(nd[:, 3] >= mo["r"] * (1 - 1e-7))
    assert np.all(nd[leaf, 3] <= mo["r"][leaf] * (1 + 3e-6))
    np.testing.assert_allclose(nd[:, 4], mo["s0"], rtol=2e-6)
    np.testing.assert_allclose(nd[:, 5], mo["s2"], rtol=2e-5)
    np.testing.assert_allclose(nd[:, 6], mo["sg"], rtol=2e-6)
    scale_m = mo["s0"][:, None]
    assert np.all(np.abs(nd[:, 7:10] - mo["M"]) <= 2e-6 * scale_m)
    scale_d = (mo["s0"] * mo["r"])[:, None]
    assert np.all(np.abs(nd[:, 10:19] - mo["D"].reshape(-1, 9)) <= 2e-5 * scale_d)
    # quadrupole by the parallel-axis shift against the particles' own sums
    qd = built["quad"]
    Qg = np.array([FF.quad_from_slots(row) for row in qd[:, :18]])
    scale_q = (mo["s0"] * mo["r"] ** 2)[:, None, None, None]
    assert np.all(np.abs(Qg - mo["Q"]) <= 5e-5 * scale_q)
    # s1, s3: exact on the leaves, bounds above
    np.testing.assert_allclose(qd[leaf, 18], mo["s1"][leaf], rtol=2e-6)
    np.testing.assert_allclose(qd[leaf, 19], mo["s3"][leaf], rtol=2e-5)
    assert np.all(qd[:, 18] >= mo["s1"] * (1 - 2e-6))
    assert np.all(qd[:, 19] >= mo["s3"] * (1 - 2e-5))
    assert np.all(qd[:, 19] <= nd[:, 3] * nd[:, 5] * (1 + 1e-5))  # s3 <= r s2


def test_lists_cover_every_pair_once(built):
    nd, info = built["nodes"], built["info"]
    first, last = nd[:, 19].astype(np.int64), nd[:, 20].astype(np.int64)
    nl = info["leaves"]
    assert info["overflow"] == 0
    assert 0 <= info["near_core_pairs"] <= info["near_pairs"]
    assert FF.list_partition_errors(first, last, nl, built["near_cnt"], built["near"],
                                    built["far_cnt"], built["far"]) == 0
    for T in range(nl):
        assert T in built["near"][T, :built["near_cnt"][T]]


def test_lists_grow_past_a_small_capacity(helix, built, monkeypatch):
    """Lists longer than their capacity are grown and built again: from an
    initial capacity of 16 entries the lists and the rates are those built
    with the default capacity, bit for bit."""
    monkeypatch.setenv("PAWAN_FF_LIST_CAP", "16")
    with capi.Solver(helix) as g:
        g.ff_set_params(0.5, 0.0, use_mlp=False)
        g.ff_set_conserve(False)
        g.ff_build()
        info = g.ff_info()
        nc, near = g.ff_export(capi.FF_NEAR_CNT), g.ff_export(capi.FF_NEAR)
        fc, far = g.ff_export(capi.FF_FAR_CNT), g.ff_export(capi.FF_FAR)
        u, w = g.solve("learned")
    _say(f"grown lists: cap_near {info['cap_near']} cap_far {info['cap_far']}")
    assert info["overflow"] == 0
    assert info["cap_near"] >= info["max_near"] > 16 and info["cap_far"] >= info["max_far"] > 16
    np.testing.assert_array_equal(nc, built["near_cnt"])
    np.testing.assert_array_equal(fc, built["far_cnt"])
    for T in range(info["leaves"]):
        np.testing.assert_array_equal(near[T, :nc[T]], built["near"][T, :nc[T]])
        np.testing.assert_array_equal(far[T, :fc[T]], built["far"][T, :fc[T]])
    np.testing.assert_array_equal(u, built["u"])
    np.testing.assert_array_equal(w, built["w"])


def test_far_nodes_pass_the_opening_test(built):
    nd, nl = built["nodes"], built["info"]["leaves"]
    for T in range(0, nl, 7):
        for q in built["far"][T, :built["far_cnt"][T]]:
            d = np.linalg.norm(nd[q, 0:3] - nd[T, 0:3])
            assert nd[q, 3] + nd[T, 3] < 0.5 * d * (1 + 1e-5)


def _oracle_rates(oracle, helix, built, targets_sorted, quad=True):
    """FP64 rates of sorted targets with the GPU's own lists: near leaves
    by the oracle's pair sums, far nodes by the expansion (to quadrupole
    order, or monopole + dipole for the network path)."""
    perm, nd = built["perm"], built["nodes"]
    sp, sa = helix.pos[perm], helix.alpha[perm]
    sr, sv = helix.radius[perm], helix.volume[perm]
    us, ws = [], []
    nl = built["info"]["leaves"]
    for t in targets_sorted:
        T = int(np.searchsorted(nd[:nl, 19], t, side="right")) - 1
        src = []
        for q in built["near"][T, :built["near_cnt"][T]]:
            src.extend(range(int(nd[q, 19]), int(nd[q, 20])))
        src = [s for s in src if s != t]
        idx = np.array([t] + src)
        sub = W.ParticleSystem(off=np.array([0, len(idx)]), pos=sp[idx], alpha=sa[idx], radius=sr[idx],
                               volume=sv[idx], nu=helix.nu, coupled=True, name="near")
        u, w = oracle.rates_targets(sub, [0])
        u, w = u[0].copy(), w[0].copy()
        for q in built["far"][T, :built["far_cnt"][T]]:
            r = (sp[t] - nd[q, 0:3])[None]
            um, Gm = FF.md_eval(r, nd[q, 7:10], nd[q, 10:19].reshape(3, 3))
            if quad:
                uq, Gq = FF.quad_eval(r, FF.quad_from_slots(built["quad"][q, :18]))
                um, Gm = um + uq, Gm + Gq
            u += um[0] / FF.FOUR_PI
            w += sa[t] @ Gm[0] / FF.FOUR_PI
        us.append(u)
        ws.append(w)
    return np.array(us), np.array(ws)


def test_analytic_far_field_matches_oracle_on_gpu_lists(oracle, helix, built):
    perm = built["perm"]
    rng = np.random.default_rng(0)
    targets = np.sort(rng.choice(helix.n, size=48, replace=False))
    ru, rw = _oracle_rates(oracle, helix, built, targets)
    gu, gw = built["u"][perm[targets]], built["w"][perm[targets]]
    assert normwise(gu, ru) < 1e-5, normwise(gu, ru)
    assert normwise(gw, rw) < 1e-5, normwise(gw, rw)


def test_near_field_with_particles_of_many_radii(oracle, helix):
    """Radii spread by +-30% (each pair then has its own core): the near
    field's general form, against the oracle on the GPU's own lists (the
    helical wake's single radius runs the form with the core factors fixed
    per launch)."""
    rng = np.random.default_rng(21)
    s = W.ParticleSystem(off=helix.off.copy(), pos=helix.pos.copy(), alpha=helix.alpha.copy(),
                         radius=helix.radius * rng.uniform(0.7, 1.3, helix.n), volume=helix.volume.copy(),
                         nu=helix.nu, coupled=True, name="helix_radii")
    with capi.Solver(s) as g:
        g.ff_set_params(0.5, 0.0, use_mlp=False)
        g.ff_set_conserve(False)
        g.ff_build()
        b = {k: g.ff_export(w) for k, w in (("perm", capi.FF_PERM), ("nodes", capi.FF_NODES),
                                            ("near_cnt", capi.FF_NEAR_CNT), ("near", capi.FF_NEAR),
                                            ("far_cnt", capi.FF_FAR_CNT), ("far", capi.FF_FAR),
                                            ("quad", capi.FF_QUAD))}
        b["info"] = g.ff_info()
        u, w = g.solve("learned")
    targets = np.sort(rng.choice(s.n, size=40, replace=False))
    ru, rw = _oracle_rates(oracle, s, b, targets)
    gu, gw = u[b["perm"][targets]], w[b["perm"][targets]]
    _say(f"many radii: e_u={normwise(gu, ru):.2e} e_w={normwise(gw, rw):.2e}")
    assert normwise(gu, ru) < 1e-5 and normwise(gw, rw) < 1e-5


def _residual_at_leaf(params, nd, T, far_ids):
    """U, G of the network summed over T's far nodes at T's centre (oracle,
    BF16 roundings emulated), 4 pi factored."""
    c = nd[T, 0:3]
    if len(far_ids) == 0:
        return np.zeros(3), np.zeros((3, 3))
    X, su, sG, Rs = [], [], [], []
    for q in far_ids:
        x, a, b, R = FF.features((c - nd[q, 0:3])[None], nd[q, 3] + nd[q, 6], nd[q, 4], nd[q, 7:10],
                                 nd[q, 10:19].reshape(3, 3), nd[q, 6])
        X.append(x[0])
        su.append(a[0])
        sG.append(b[0])
        Rs.append(R[0])
    Y = FF.mlp_forward(params, np.array(X, dtype=np.float32), bf16_emulate=True)
    U, G = np.zeros(3), np.zeros((3, 3))
    for k, R in enumerate(Rs):  # back from each node's frame
        U += su[k] * (R.T @ Y[k, :3])
        G += sG[k] * (R.T @ Y[k, 3:].reshape(3, 3) @ R)
    return U, G


def test_learned_far_field_matches_oracle_on_gpu_lists(oracle, helix, built):
    """The network's residual at the leaf centres, carried to the targets
    by its gradient output, on top of the analytic far field."""
    params = _rand_params(12, scale=0.5)
    perm, nd = built["perm"], built["nodes"]
    with capi.Solver(helix) as g:
        g.ff_set_weights(params)
        g.ff_set_params(0.5, 0.0, use_mlp=True)
        g.ff_set_conserve(False)
        u, w = g.solve("learned")
    rng = np.random.default_rng(1)
    targets = np.sort(rng.choice(helix.n, size=32, replace=False))
    ru, rw = _oracle_rates(oracle, helix, built, targets, quad=False)
    sp, sa = helix.pos[perm], helix.alpha[perm]
    du = []
    nl = built["info"]["leaves"]
    for k, t in enumerate(targets):
        T = int(np.searchsorted(nd[:nl, 19], t, side="right")) - 1
        U, G = _residual_at_leaf(params, nd, T, built["far"][T, :built["far_cnt"][T]])
        du.append(np.linalg.norm(U) / np.linalg.norm(ru[k] * FF.FOUR_PI))
        ru[k] += (U + G @ (sp[t] - nd[T, 0:3])) / FF.FOUR_PI
        rw[k] += sa[t] @ G / FF.FOUR_PI
    _say(f"network residual / velocity: median {np.median(du):.2e}, max {np.max(du):.2e}")
    assert np.median(du) > 1e-4  # the residual is large enough to be checked
    gu, gw = u[perm[targets]], w[perm[targets]]
    assert normwise(gu, ru) < 2e-5, normwise(gu, ru)
    assert normwise(gw, rw) < 2e-5, normwise(gw, rw)


def test_conservation_removes_the_mean_vorticity_rate(helix, built):
    """With conservation on (the default), sum dalpha vanishes to rounding,
    as in the direct sum, and each rate differs from the raw one by the
    same vector, the mean."""
    with capi.Solver(helix) as g:
        assert g.ff_conserve()
        g.ff_set_params(0.5, 0.0, use_mlp=False)
        u, w = g.solve("learned")
        ud, wd = g.solve("fp64")
    scale = np.abs(built["w"]).sum(axis=0)
    raw_sum = built["w"].sum(axis=0)
    _say(f"sum dalpha raw {raw_sum}, conserved {w.sum(axis=0)}, direct {wd.sum(axis=0)}")
    assert np.all(np.abs(w.sum(axis=0)) <= 1e-12 * scale)
    np.testing.assert_array_equal(u, built["u"])
    shift = built["w"] - w
    np.testing.assert_allclose(shift, np.broadcast_to(raw_sum / helix.n, shift.shape), rtol=1e-9,
                               atol=1e-15 * np.abs(built["w"]).max())


def test_error_against_direct_sum_falls_with_theta(helix):
    errs = {}
    with capi.Solver(helix) as g:
        ud, wd = g.solve("fp64")
        for theta in (0.6, 0.45, 0.3):
            g.ff_set_params(theta, 0.0, use_mlp=False)
            u, w = g.solve("learned")
            errs[theta] = (normwise(u, ud), normwise(w, wd))
            _say(f"theta {theta}: u {errs[theta][0]:.2e} w {errs[theta][1]:.2e} {g.ff_info()}")
    assert errs[0.6][0] < 3e-2
    assert errs[0.3][0] < errs[0.45][0] < errs[0.6][0]
    assert errs[0.3][1] < errs[0.6][1]


def test_error_tolerance_opens_nodes(helix):
    """The error test tightens the lists: a smaller tolerance gives more
    near pairs and a smaller error."""
    with capi.Solver(helix) as g:
        ud, _ = g.solve("fp64")
        urms = float(np.sqrt((ud * ud).sum(axis=1).mean()))
        res = []
        for tol in (0.0, 1e-3, 1e-4):
            g.ff_set_params(0.5, tol * urms, use_mlp=False)
            u, _ = g.solve("learned")
            res.append((normwise(u, ud), g.ff_info()["near_pairs"]))
            _say(f"tol {tol}: err {res[-1][0]:.2e} near pairs {res[-1][1]}")
    assert res[1][0] <= res[0][0] and res[2][0] < res[1][0]
    assert res[2][1] > res[1][1] >= res[0][1]


def test_samples_match_oracle(helix, built):
    with capi.Solver(helix) as g:
        g.ff_set_params(0.5, 0.0, use_mlp=False)
        g.ff_sample(40, seed=9)
        X = g.ff_export(capi.FF_SAMPLES_X)
        Y = g.ff_export(capi.FF_SAMPLES_Y)
        A = g.ff_export(capi.FF_SAMPLES_AUX)
    perm, nd = built["perm"], built["nodes"]
    sp, sa, sr = helix.pos[perm], helix.alpha[perm], helix.radius[perm]
    for k in range(len(X)):
        q, t = int(A[k, 7]), int(A[k, 8])
        a, b = int(nd[q, 19]), int(nd[q, 20])
        assert not (a <= t < b)
        ue, Ge = FF.exact_node(sp[t], sr[t], sp[a:b], sa[a:b], sr[a:b])
        r = sp[t] - nd[q, 0:3]
        um, Gm = FF.md_eval(r[None], nd[q, 7:10], nd[q, 10:19].reshape(3, 3))
        uq, _ = FF.quad_eval(r[None], FF.quad_from_slots(built["quad"][q, :18]))
        # side data: the error of the expansion to quadrupole order, and s3
        eq = np.linalg.norm(ue - um[0] - uq[0])
        assert abs(A[k, 10] - eq) <= 1e-5 * np.linalg.norm(um[0]) + 1e-3 * eq, k
        np.testing.assert_allclose(A[k, 11], built["quad"][q, 19], rtol=1e-6)
        xo, su, sG, R = FF.features(r[None], nd[q, 3] + nd[q, 6], nd[q, 4], nd[q, 7:10],
                                    nd[q, 10:19].reshape(3, 3), nd[q, 6])
        np.testing.assert_allclose(X[k], xo[0], rtol=1e-5, atol=5e-6)  # FP32 rotations: a few ulp of O(1)
        # residual in the node frame against the FP64 residual, relative to
        # the expansion's size (the GPU expansion is FP32)
        R = R[0]
        res_u = Y[k, :3] * su[0]
        res_G = Y[k, 3:].reshape(3, 3) * sG[0]
        assert np.linalg.norm(res_u - R @ (ue - um[0])) <= 1e-5 * np.linalg.norm(um[0]), k
        assert np.linalg.norm(res_G - R @ (Ge - Gm[0]) @ R.T) <= 1e-5 * np.linalg.norm(Gm[0]), k
        assert A[k, 4] == pytest.approx(np.linalg.norm(ue - um[0]), rel=1e-3, abs=1e-6 * np.linalg.norm(um[0]))


def _rand_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    p = np.zeros(FF.N_PARAMS, dtype=np.float32)
    for name, (o, size) in FF.OFF.items():
        fan_in = {"W1": 16, "W2": 128, "W3": 128, "W4": 128}.get(name, 128)
        p[o:o + size] = rng.normal(size=size) * scale / np.sqrt(fan_in) if name[0] == "W" else rng.normal(size=size) * 0.1
    return p


def test_mlp_inference_matches_bf16_oracle(helix):
    """The matrix-core network against the oracle with the same BF16
    roundings (inputs, weights, activations); the FP64 network without them
    differs visibly, so the comparison is sensitive to the roundings, and a
    wrong operand layout would be an O(1) error."""
    params = _rand_params(4)
    rng = np.random.default_rng(5)
    X = rng.normal(size=(1000, 16)).astype(np.float32)
    with capi.Solver(helix) as g:
        g.ff_set_weights(params)
        Yg = g.ff_mlp_forward(X)
        np.testing.assert_array_equal(g.ff_get_weights(), params)
    Ye = FF.mlp_forward(params, X, bf16_emulate=True)
    Yf = FF.mlp_forward(params, X, bf16_emulate=False)
    e_emu, e_fp64 = normwise(Yg, Ye), normwise(Yg, Yf)
    _say(f"mlp forward: vs bf16 oracle {e_emu:.2e}, vs fp64 oracle {e_fp64:.2e}")
    assert e_emu < 3e-3
    assert e_emu < 0.5 * e_fp64


def test_mlp_training_gradient_matches_oracle(helix):
    params = _rand_params(6, scale=0.8)
    rng = np.random.default_rng(7)
    n = 100  # not a multiple of the 32-row block
    X = rng.normal(size=(n, 16)).astype(np.float32)
    Y = rng.normal(size=(n, 12)).astype(np.float32)
    with capi.Solver(helix) as g:
        g.ff_set_weights(params)
        grad, loss = g.ff_mlp_grad(X, Y)
    lo, go = FF.mlp_loss_grad(params.astype(np.float64), X.astype(np.float64), Y.astype(np.float64))
    assert loss == pytest.approx(lo, rel=1e-5)

    def rel(a, b):
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))

    assert rel(grad, go) < 1e-5, rel(grad, go)
    for name, (o, size) in FF.OFF.items():
        assert rel(grad[o:o + size], go[o:o + size]) < 1e-4, (name, rel(grad[o:o + size], go[o:o + size]))


def test_learned_rk4_step_runs_and_tracks_direct(helix):
    with capi.Solver(helix) as a, capi.Solver(helix) as b:
        a.ff_set_params(0.4, 0.0, use_mlp=False)
        a.step("rk4", 1e-3, "learned", 2)
        b.step("rk4", 1e-3, "fp32c", 2)
        pa, aa = a.state()
        pb, ab = b.state()
    disp = pb - helix.pos
    assert np.all(np.isfinite(pa)) and np.all(np.isfinite(aa))
    assert normwise(pa - helix.pos, disp) < 3e-2
    assert normwise(aa, ab) < 1e-3


def test_learned_requires_weights_and_coupling(helix):
    with capi.Solver(helix) as g:
        g.ff_set_params(0.5, 0.0, use_mlp=True)
        with pytest.raises(capi.PawanError) as e:
            g.solve("learned")
        assert e.value.code == capi.PW_ERR_STATE
    s = W.batched_rings(nwakes=3, nmin=300, nmax=600, seed=1)
    with capi.Solver(s) as g:
        g.ff_set_params(0.5, 0.0, use_mlp=False)
        with pytest.raises(capi.PawanError) as e:
            g.solve("learned")
        assert e.value.code == capi.PW_ERR_UNSUPPORTED
