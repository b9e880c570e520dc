"""CPU pins of the selection oracle (oracle/mlp_oracle.py): the analytic
gradient against central finite differences, Adam's first step against its
closed form, the Lagrangian order selection against exhaustive search, the
level features against direct counts over the oracle tree, and the model file
round trip."""
import math
import os

import numpy as np
import pytest

from oracle import mlp_oracle as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _problem(seed=0, rows_per=(3, 5, 7, 4, 6, 2)):
    rng = np.random.Generator(np.random.PCG64(seed))
    params = rng.normal(0, 0.3, M.NP)
    seg = np.concatenate([[0], np.cumsum(rows_per)]).astype(np.int64)
    X = rng.normal(0, 1, (seg[-1], M.F))
    y = rng.normal(0, 2, (len(rows_per), 2))
    return params, X, seg, y


def test_gradient_matches_finite_differences():
    params, X, seg, y = _problem()
    loss, g = M.loss_and_grad(params, X, seg, y, 0.7)
    assert abs(loss - M.loss_only(params, X, seg, y, 0.7)) < 1e-12
    rng = np.random.Generator(np.random.PCG64(1))
    # every block is probed (indices drawn per block)
    idx = []
    for k, (o, s) in M.OFF.items():
        n = int(np.prod(s))
        idx += list(o + rng.choice(n, min(n, 12), replace=False))
    h = 1e-6
    for k in idx:
        pp, pm = params.copy(), params.copy()
        pp[k] += h
        pm[k] -= h
        fd = (M.loss_only(pp, X, seg, y, 0.7) - M.loss_only(pm, X, seg, y, 0.7)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])


def test_gradient_of_a_planted_sign_error_fails():
    """A wrong sign in the tanh derivative is caught by the same check."""
    params, X, seg, y = _problem(2)
    _, g = M.loss_and_grad(params, X, seg, y, 1.0)
    P = M.unpack(params)
    out, (Xd, h1, h2) = M.forward(params, X)
    # recompute dW1 with (1 + h^2) instead of (1 - h^2)
    S = len(seg) - 1
    dout = np.zeros_like(out)
    for s in range(S):
        a, b = seg[s], seg[s + 1]
        le, lt = M._lse(out[a:b, 0]), M._lse(out[a:b, 1])
        dout[a:b, 0] = 2 * (le - y[s, 0]) * np.exp(out[a:b, 0] - le) / S
        dout[a:b, 1] = 2 * (lt - y[s, 1]) * np.exp(out[a:b, 1] - lt) / S
    dh2 = (dout @ P["W3"].T) * (1 - h2 * h2)
    dh1_bad = (dh2 @ P["W2"].T) * (1 + h1 * h1)
    bad = (Xd.T @ dh1_bad).ravel()
    o = M.OFF["W1"][0]
    assert np.max(np.abs(bad - g[o:o + bad.size])) > 1e-3


def test_adam_first_step_is_lr_times_sign():
    g = np.array([0.3, -2.0, 1e-3, -5e-4])
    p, m, v = M.adam_step(np.zeros(4), g, np.zeros(4), np.zeros(4), 1, 0.01, eps=0.0)
    assert np.allclose(p, -0.01 * np.sign(g), rtol=0, atol=1e-15)
    # second step with the same gradient: still lr * sign (bias correction exact)
    p2, _, _ = M.adam_step(p, g, m, v, 2, 0.01, eps=0.0)
    assert np.allclose(p2 - p, -0.01 * np.sign(g), atol=1e-15)


def _tables(seed, nrow=5, npv=6, convex=True):
    rng = np.random.Generator(np.random.PCG64(seed))
    e = np.zeros((nrow, npv))
    t = np.zeros((nrow, npv))
    for L in range(nrow):
        a, b = rng.uniform(-22, -14), rng.uniform(0.8, 2.5)
        c, d = rng.uniform(-3, 0), rng.uniform(0.05, 0.3)
        pp = np.arange(npv)
        e[L] = a - b * pp + (0 if convex else rng.normal(0, 0.5, npv))
        t[L] = c + d * pp + (0 if convex else rng.normal(0, 0.2, npv))
    e[:2, 1:] = np.nan
    t[:2, 1:] = np.nan
    return e, t


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("convex", [True, False])
def test_selection_against_exhaustive_search(seed, convex):
    e, t = _tables(seed, convex=convex)
    target = 3e-9
    orders, src, perr, pms = M.select_tables(e, t, 4, target, 12)
    best_t, best_c, _ = M.brute_tables(e, t, target)
    if best_c is None:
        assert src == 2
        return
    assert src == 1
    ch = [o - 4 for o in orders[2:len(e)]]
    err2 = sum(math.exp(e[L, 0]) for L in range(2)) + sum(math.exp(e[L, c]) for L, c in
                                                        zip(range(2, len(e)), ch))
    tim = sum(math.exp(t[L, 0]) for L in range(2)) + sum(math.exp(t[L, c]) for L, c in
                                                       zip(range(2, len(e)), ch))
    assert err2 <= target ** 2 * (1 + 1e-9)         # feasible
    assert tim >= best_t * (1 - 1e-12)               # not better than the optimum
    assert tim <= best_t * 1.3                       # Lagrangian gap stays small
    assert abs(math.sqrt(err2) - perr) < 1e-9 * perr and abs(tim - pms) < 1e-9 * pms


def test_selection_falls_back_when_infeasible_or_out_of_range():
    e, t = _tables(0)
    orders, src, _, _ = M.select_tables(e, t, 4, 1e-30, 12)
    assert src == 2 and set(orders) == {12}
    orders, src, _, _ = M.select_tables(e, t, 4, 1.0, 12, ok=False)
    assert src == 2 and set(orders) == {12}


def test_level_features_against_direct_counts(oracle):
    from lsmps_vpm_v2_amd import workloads as W
    f = W.random_clusters(4000, seed=5)
    o = oracle.opts(max_par=40, max_src=40)
    tree = oracle.Tree(f["x"], f["y"], f["active"], o)
    feats, raw = M.level_features(tree, f["q"], 40)
    cells = tree.cells()
    lid, poff, pidx = tree.leaves()
    lev_of = dict(zip(cells["id"].tolist(), cells["level"].tolist()))
    nlev = len(raw)
    aq = np.abs(f["q"])
    for L in range(nlev):
        under = [k for k, c in enumerate(lid) if lev_of[int(c)] >= L]
        parts = np.concatenate([pidx[poff[k]:poff[k + 1]] for k in under])
        assert raw[L, 1] == len(parts)                      # particles under level L
        assert abs(raw[L, 2] - aq[parts].sum()) < 1e-12 * aq.sum()
        assert raw[L, 3] == aq[parts].max()
        assert raw[L, 0] == np.sum(cells["level"] == L)
    assert raw[0, 1] == f["n"] and raw[0, 0] == 1
    assert np.allclose(feats[:, 0], np.arange(nlev))
    R, _ = tree.geometry()
    assert np.allclose(feats[:, 3], math.log2(R) - np.arange(nlev))
    assert np.all(feats[:, 4] <= 0.0) and feats[0, 4] == 0.0
    assert np.allclose(feats[:, 7], math.log2(40))


def test_model_file_round_trip(tmp_path):
    from lsmps_vpm_v2_amd import selector as S
    rng = np.random.Generator(np.random.PCG64(3))
    m = S.Model(params=rng.normal(0, 1, S.NP).astype(np.float32),
                mu=rng.normal(0, 1, 9).astype(np.float32),
                sd=rng.uniform(0.5, 2, 9).astype(np.float32),
                lo=-np.ones(9, np.float32), hi=np.ones(9, np.float32), p_min=5, p_max=14,
                thresholds=[24, 48, 96], target_err=2e-4, margin=0.1)
    path = str(tmp_path / "m.bin")
    m.save(path)
    r = S.Model.load(path)
    assert np.array_equal(r.params, m.params) and np.array_equal(r.mu, m.mu)
    assert np.array_equal(r.sd, m.sd) and (r.p_min, r.p_max) == (5, 14)
    assert r.thresholds == [24, 48, 96] and abs(r.target_err - 2e-4) < 1e-10
    st = r.to_struct()
    assert st.n_params == 4930 and st.n_thresholds == 3 and st.thresholds[2] == 96
