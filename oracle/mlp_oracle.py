"""CPU oracle (test infrastructure only) of the learned order selection.

Float64 numpy restatement of what lsmps_vpm_v2_amd/csrc/select.hip computes:
the per-level feature rows of a tree, the 9 -> 64 tanh -> 64 tanh -> 2 MLP,
the per-solve loss (log-sum-exp of the level outputs against ln err^2 and
ln ms), its gradient, one Adam step, and the Lagrangian order selection.
The reference has no counterpart (its order is one global P_max,
global.cpp:311).  Pinning: the gradient against central finite differences,
the selection against exhaustive search on small trees, the features against
direct counts over the oracle tree's particle lists (tests/test_oracle_mlp.py).

Only tests/, ``__graft_entry__.smoke`` and bench.py's cpu_baseline may use it.
"""
from __future__ import annotations

import math

import numpy as np

F, H, NP = 9, 64, 4930
OFF = {"W1": (0, (F, H)), "b1": (F * H, (H,)), "W2": (F * H + H, (H, H)),
       "b2": (F * H + H + H * H, (H,)), "W3": (F * H + 2 * H + H * H, (H, 2)),
       "b3": (F * H + 4 * H + H * H, (2,))}


def unpack(params):
    p = np.asarray(params, np.float64)
    return {k: p[o:o + int(np.prod(s))].reshape(s) for k, (o, s) in OFF.items()}


def pack(blocks) -> np.ndarray:
    out = np.zeros(NP, np.float64)
    for k, (o, s) in OFF.items():
        out[o:o + int(np.prod(s))] = np.asarray(blocks[k], np.float64).ravel()
    return out


# ---------------------------------------------------------------------------
# features
# ---------------------------------------------------------------------------
def level_features(tree, q, threshold):
    """(feats[nlev, 9], raw[nlev, 5]) of an oracle ``Tree`` with sources q:
    per level L the cells, particles under the level's cells, sum |q| and
    max |q| of those particles, and leaves; features as in select.hip
    (order column 0)."""
    cells = tree.cells()
    lid, poff, pidx = tree.leaves()
    R, _ = tree.geometry()
    lev_of = dict(zip(cells["id"].tolist(), cells["level"].tolist()))
    nlev = int(cells["level"].max()) + 1
    raw = np.zeros((nlev, 5))
    for L in range(nlev):
        m = cells["level"] == L
        raw[L, 0] = m.sum()
        raw[L, 1] = cells["par_num"][m].sum()
        raw[L, 4] = (cells["is_leaf"][m] != 0).sum()
    aq = np.abs(np.asarray(q, np.float64))
    for k, cid in enumerate(lid):
        Lf = lev_of[int(cid)]
        idx = pidx[poff[k]:poff[k + 1]]
        s, mx = aq[idx].sum(), (aq[idx].max() if len(idx) else 0.0)
        raw[:Lf + 1, 2] += s
        raw[:Lf + 1, 3] = np.maximum(raw[:Lf + 1, 3], mx)
    feats = np.zeros((nlev, F))
    sum0 = raw[0, 2] if raw[0, 2] > 0 else 1.0
    n0 = raw[0, 1] if raw[0, 1] > 0 else 1.0
    for L in range(nlev):
        c = raw[L, 0]
        if c <= 0:
            continue
        feats[L] = [L, math.log2(c), math.log2(max(raw[L, 1], 1.0) / c), math.log2(R) - L,
                    math.log10(max(raw[L, 2] / sum0, 1e-12)),
                    math.log10(max(raw[L, 3] * n0 / sum0, 1e-12)), raw[L, 4] / c,
                    math.log2(threshold), 0.0]
    return feats, raw


# ---------------------------------------------------------------------------
# network, loss, gradient, Adam
# ---------------------------------------------------------------------------
def forward(params, X):
    P = unpack(params)
    X = np.asarray(X, np.float64)
    h1 = np.tanh(X @ P["W1"] + P["b1"])
    h2 = np.tanh(h1 @ P["W2"] + P["b2"])
    out = h2 @ P["W3"] + P["b3"]
    return out, (X, h1, h2)


def _lse(v):
    m = v.max()
    return m + math.log(np.exp(v - m).sum())


def loss_and_grad(params, X, seg, y, w_time):
    """loss = mean_s (LSE_s(e) - y[s, 0])^2 + w (LSE_s(t) - y[s, 1])^2 and its
    gradient with respect to the flat parameter vector."""
    P = unpack(params)
    out, (X, h1, h2) = forward(params, X)
    y = np.asarray(y, np.float64).reshape(-1, 2)
    S = len(seg) - 1
    dout = np.zeros_like(out)
    loss = 0.0
    for s in range(S):
        a, b = seg[s], seg[s + 1]
        le, lt = _lse(out[a:b, 0]), _lse(out[a:b, 1])
        de, dt = le - y[s, 0], lt - y[s, 1]
        loss += de * de + w_time * dt * dt
        dout[a:b, 0] = 2 * de * np.exp(out[a:b, 0] - le) / S
        dout[a:b, 1] = 2 * w_time * dt * np.exp(out[a:b, 1] - lt) / S
    loss /= S
    g = {"W3": h2.T @ dout, "b3": dout.sum(0)}
    dh2 = (dout @ P["W3"].T) * (1 - h2 * h2)
    g["W2"] = h1.T @ dh2
    g["b2"] = dh2.sum(0)
    dh1 = (dh2 @ P["W2"].T) * (1 - h1 * h1)
    g["W1"] = X.T @ dh1
    g["b1"] = dh1.sum(0)
    return loss, pack(g)


def loss_only(params, X, seg, y, w_time):
    out, _ = forward(params, X)
    y = np.asarray(y, np.float64).reshape(-1, 2)
    S = len(seg) - 1
    tot = 0.0
    for s in range(S):
        a, b = seg[s], seg[s + 1]
        tot += (_lse(out[a:b, 0]) - y[s, 0]) ** 2 + w_time * (_lse(out[a:b, 1]) - y[s, 1]) ** 2
    return tot / S


def adam_step(params, grad, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    c1, c2 = 1 - b1 ** step, 1 - b2 ** step
    return params - lr * (m / c1) / (np.sqrt(v / c2) + eps), m, v


def cosine_lr(k, steps, lr, lr_min):
    return lr_min + 0.5 * (lr - lr_min) * (1 + math.cos(math.pi * k / steps))


# ---------------------------------------------------------------------------
# selection
# ---------------------------------------------------------------------------
def predictions(model, feats):
    """e[L, i], t[L, i] for orders p_min + i (levels 0, 1: order 0 in column 0)."""
    nrow = len(feats)
    npv = model["p_max"] - model["p_min"] + 1
    e = np.full((nrow, npv), np.nan)
    t = np.full((nrow, npv), np.nan)
    for L in range(nrow):
        for i in range(npv if L >= 2 else 1):
            x = np.array(feats[L], np.float64)
            x[F - 1] = 0.0 if L < 2 else model["p_min"] + i
            xn = (x - model["mu"]) / model["sd"]
            o, _ = forward(model["params"], xn[None, :])
            e[L, i], t[L, i] = o[0]
    return e, t


def in_range(model, feats):
    lo, hi = np.asarray(model["lo"], np.float64), np.asarray(model["hi"], np.float64)
    mg = model["margin"]
    span = hi - lo
    f = np.asarray(feats, np.float64)[:, :F - 1]
    return bool(np.all(f >= lo[:F - 1] - mg * span[:F - 1]) and
                np.all(f <= hi[:F - 1] + mg * span[:F - 1]))


SEL_AIM = 0.7  # select.hip aims at 0.7 x the target (the judge enforces the target)


def select(model, feats, fixed_p):
    """(orders[18], source, pred_err, pred_ms) by the Lagrangian bisection of
    select.hip (see select_tables) at 0.7 x the model's target."""
    e, t = predictions(model, feats)
    return select_tables(e, t, model["p_min"], SEL_AIM * model["target_err"], fixed_p,
                         in_range(model, feats))


def select_tables(e, t, p_min, target_err, fixed_p, ok=True):
    """p_L(lam) = argmin_p exp(t[L, p]) + exp(lam + e[L, p]); the smallest lam
    in [-10, 80] (30 halvings) with sum_L exp(e) <= target^2; rows 0 and 1
    contribute column 0 as constants."""
    nrow = len(e)
    tgt = 2 * math.log(target_err)
    c_err = sum(math.exp(e[L, 0]) for L in range(min(2, nrow)))
    c_time = sum(math.exp(t[L, 0]) for L in range(min(2, nrow)))

    def pick(lam):
        ch, err2, tim = {}, c_err, c_time
        for L in range(2, nrow):
            v = np.exp(t[L]) + np.exp(lam + e[L])
            b = int(np.argmin(v))
            ch[L] = b
            err2 += math.exp(e[L, b])
            tim += math.exp(t[L, b])
        return ch, err2, tim

    lo, hi = -10.0, 80.0
    ch, err2, tim = pick(hi)
    if not ok or not (math.log(err2) <= tgt):
        return [fixed_p] * 18, 2, -1.0, -1.0
    for _ in range(30):
        mid = 0.5 * (lo + hi)
        _, e2, _ = pick(mid)
        if math.log(e2) <= tgt:
            hi = mid
        else:
            lo = mid
    ch, err2, tim = pick(hi)
    orders = [fixed_p] * 18
    for L in range(2, nrow):
        orders[L] = p_min + ch[L]
    if nrow > 2:
        orders[0] = orders[1] = orders[2]
    return orders, 1, math.sqrt(err2), tim


def brute_tables(e, t, target_err):
    """Exhaustive (time, choice, err^2) minimum of predicted time subject to
    predicted error <= target over every order vector (small tables only)."""
    import itertools
    nrow = len(e)
    c_err = sum(math.exp(e[L, 0]) for L in range(min(2, nrow)))
    c_time = sum(math.exp(t[L, 0]) for L in range(min(2, nrow)))
    best = (math.inf, None, None)
    for combo in itertools.product(range(e.shape[1]), repeat=max(nrow - 2, 0)):
        err2 = c_err + sum(math.exp(e[L, c]) for L, c in zip(range(2, nrow), combo))
        tim = c_time + sum(math.exp(t[L, c]) for L, c in zip(range(2, nrow), combo))
        if err2 <= target_err ** 2 and tim < best[0]:
            best = (tim, combo, err2)
    return best
