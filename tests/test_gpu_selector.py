"""GPU tests of the learned order selection: the MLP kernels against the
float64 oracle (forward, loss, gradient, Adam, a training run), the device
level features against the oracle tree, the device selection against the
oracle's selection, the fixed-order fallback, and the gain of the trained
model over the fixed order p = 12."""
import math
import os

import numpy as np
import pytest

from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu


def _rand_params(seed):
    from lsmps_vpm_v2_amd import selector as S
    rng = np.random.Generator(np.random.PCG64(seed))
    p = S.init_params(seed, (-3.0, 0.5)).astype(np.float64)
    p += rng.normal(0, 0.05, p.size)
    return p.astype(np.float32)


def _segments(seed, nseg=150):
    rng = np.random.Generator(np.random.PCG64(seed))
    sizes = rng.integers(1, 18, nseg)
    seg = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    X = rng.normal(0, 1, (seg[-1], M.F)).astype(np.float32)
    y = rng.normal(0, 1.5, (nseg, 2)).astype(np.float32)
    return X, seg, y


def test_mlp_forward_matches_oracle(gpu):
    import torch
    from lsmps_vpm_v2_amd import FmmSolver
    s = FmmSolver()
    params = _rand_params(1)
    X, _, _ = _segments(1)
    out = s.mlp_forward(torch.from_numpy(params).cuda(), torch.from_numpy(X).cuda()).cpu().numpy()
    ref, _ = M.forward(params.astype(np.float64), X.astype(np.float64))
    print(f"forward: rows {len(X)}, max |diff| {np.max(np.abs(out - ref)):.2e}")
    assert np.allclose(out, ref, rtol=1e-5, atol=2e-5)


def test_mlp_gradient_matches_oracle(gpu):
    import torch
    from lsmps_vpm_v2_amd import FmmSolver
    s = FmmSolver()
    params = _rand_params(2)
    X, seg, y = _segments(2)
    assert len(X) > 4 * 256  # several weight-gradient chunks, the last one partial
    loss, g = s.mlp_loss_grad(torch.from_numpy(params).cuda(), torch.from_numpy(X).cuda(),
                              torch.from_numpy(seg).cuda(), torch.from_numpy(y.ravel()).cuda(),
                              0.5)
    g = g.cpu().numpy().astype(np.float64)
    lref, gref = M.loss_and_grad(params.astype(np.float64), X, seg, y, 0.5)
    rel = np.linalg.norm(g - gref) / np.linalg.norm(gref)
    print(f"loss {loss:.6f} vs {lref:.6f}; gradient rel-L2 {rel:.2e}")
    assert abs(loss - lref) <= 1e-5 * abs(lref)
    assert rel <= 1e-4
    for k, (o, shp) in M.OFF.items():  # every parameter block, elementwise
        n = int(np.prod(shp))
        scale = np.max(np.abs(gref[o:o + n])) + 1e-12
        assert np.max(np.abs(g[o:o + n] - gref[o:o + n])) <= 2e-4 * scale, k


def test_adam_step_and_training_match_oracle(gpu):
    import torch
    from lsmps_vpm_v2_amd import FmmSolver
    s = FmmSolver()
    params = _rand_params(3)
    X, seg, y = _segments(3, nseg=60)
    P = torch.from_numpy(params).cuda()
    Mm, V = torch.zeros_like(P), torch.zeros_like(P)
    Xd, Sd, Yd = (torch.from_numpy(a).cuda() for a in (X, seg, y.ravel()))
    h = s.mlp_train(P, Mm, V, Xd, Sd, Yd, 1.0, steps=1, lr=1e-3, lr_min=1e-3)
    l0, g0 = M.loss_and_grad(params.astype(np.float64), X, seg, y, 1.0)
    p1, _, _ = M.adam_step(params.astype(np.float64), g0, 0 * g0, 0 * g0, 1, 1e-3)
    assert abs(h[0] - l0) <= 1e-5 * abs(l0)
    assert np.allclose(P.cpu().numpy(), p1, rtol=0, atol=2e-6)
    # continue: 300 steps with a cosine schedule; the float64 oracle follows
    steps = 300
    hist = s.mlp_train(P, Mm, V, Xd, Sd, Yd, 1.0, steps=steps, step0=1, lr=3e-3, lr_min=1e-4)
    pr, mr, vr = p1, 0.1 * g0, 0.001 * g0 * g0
    ref = []
    for k in range(steps):
        lk, gk = M.loss_and_grad(pr, X, seg, y, 1.0)
        ref.append(lk)
        pr, mr, vr = M.adam_step(pr, gk, mr, vr, k + 2, M.cosine_lr(k, steps, 3e-3, 1e-4))
    ref = np.asarray(ref)
    print(f"training: loss {hist[0]:.4f} -> {hist[-1]:.4f} (oracle {ref[0]:.4f} -> {ref[-1]:.4f})")
    assert hist[-1] < 0.5 * hist[0]
    assert abs(hist[0] - ref[0]) <= 1e-4 * abs(ref[0])
    # Adam normalises each coordinate, so FP32 rounding of near-zero gradients
    # moves the trajectory slightly; the losses stay within a few percent
    for k in (50, 150, steps - 1):
        assert abs(hist[k] - ref[k]) <= 5e-2 * abs(ref[k]) + 1e-3, (k, hist[k], ref[k])


@pytest.mark.parametrize("thr", [24, 64])
def test_level_features_match_oracle_tree(gpu, oracle, thr):
    import torch
    from lsmps_vpm_v2_amd import FmmSolver, workloads as W
    f = W.random_clusters(30000, seed=9, active_frac=0.8)
    s = FmmSolver(p=8, max_par=thr, max_src=thr, precision="mixed", binom_mode=0, list4_mode=2)
    dev = [torch.from_numpy(f[k]).cuda() for k in ("x", "y", "s", "omega", "active")]
    s.velocity_device(*dev)
    feats, raw = s.level_features()
    tree = oracle.Tree(f["x"], f["y"], f["active"], oracle.opts(max_par=thr, max_src=thr))
    rf, rr = M.level_features(tree, f["q"], thr)
    assert feats.shape == rf.shape
    assert np.array_equal(raw[:, [0, 1, 4]], rr[:, [0, 1, 4]])      # counts exact
    assert np.allclose(raw[:, 2], rr[:, 2], rtol=1e-5)              # sum |q| (FP32)
    assert np.allclose(raw[:, 3], rr[:, 3], rtol=1e-6)              # max |q|
    assert np.allclose(feats, rf, atol=2e-5)


def _model_for(feats, seed, target_scale):
    """A random model whose normalisation covers feats; target between the
    smallest and largest predicted error of the tree."""
    params = _rand_params(seed).astype(np.float64)
    X = np.array(feats, np.float64)
    mu = X.mean(axis=0)
    sd = X.std(axis=0) + 0.5
    mu[-1], sd[-1] = 10.0, 4.0
    lo = X.min(axis=0) - 1.0
    hi = X.max(axis=0) + 1.0
    lo[-1], hi[-1] = 4, 16
    m = dict(params=params, mu=mu, sd=sd, lo=lo, hi=hi, margin=0.25, p_min=4, p_max=16,
             target_err=1.0)
    e, t = M.predictions(m, feats)
    lo2 = sum(np.nanmin(np.exp(e[L])) for L in range(len(e)))
    hi2 = sum(np.nanmax(np.exp(e[L])) for L in range(len(e)))
    m["target_err"] = math.sqrt(lo2 + target_scale * (hi2 - lo2)) / M.SEL_AIM
    return m


def _install(solver, m, thresholds):
    from lsmps_vpm_v2_amd import selector as S
    solver.set_selector(S.Model(params=m["params"].astype(np.float32),
                                mu=m["mu"].astype(np.float32), sd=m["sd"].astype(np.float32),
                                lo=m["lo"].astype(np.float32), hi=m["hi"].astype(np.float32),
                                p_min=m["p_min"], p_max=m["p_max"], thresholds=thresholds,
                                target_err=m["target_err"], margin=m["margin"]))


@pytest.mark.parametrize("scale", [0.02, 0.3])
def test_device_selection_matches_oracle(gpu, scale):
    import torch
    from lsmps_vpm_v2_amd import FmmSolver, workloads as W
    f = W.random_clusters(40000, seed=4)
    dev = [torch.from_numpy(f[k]).cuda() for k in ("x", "y", "s", "omega", "active")]
    s = FmmSolver(p=12, max_par=48, max_src=48, precision="mixed", binom_mode=0, list4_mode=2)
    s.velocity_device(*dev)
    feats, _ = s.level_features()
    m = _model_for(feats.astype(np.float64), 5, scale)
    _install(s, m, [48])
    s.set_opts(order_select=1, judge_every=-1)  # the selection itself, not the judge
    s.velocity_device(*dev)
    sel = s.last_selection()
    orders, src, perr, pms = M.select(m, feats.astype(np.float64), 12)
    print(f"device {sel['source']} {sel['p'][:len(feats)]} / oracle {src} {orders[:len(feats)]}")
    assert sel["source"] == src == 1 and sel["threshold"] == 48
    e, t = M.predictions(m, feats.astype(np.float64))
    nrow = len(feats)

    def cost(ords):
        err2 = sum(math.exp(e[L, 0]) for L in range(2))
        tim = sum(math.exp(t[L, 0]) for L in range(2))
        for L in range(2, nrow):
            err2 += math.exp(e[L, ords[L] - 4])
            tim += math.exp(t[L, ords[L] - 4])
        return err2, tim
    if list(sel["p"][2:nrow]) != list(orders[2:nrow]):
        e2, tm = cost(sel["p"])
        _, tref = cost(orders)
        assert e2 <= (M.SEL_AIM * m["target_err"]) ** 2 * (1 + 1e-3) and tm <= tref * (1 + 1e-3)
    assert abs(sel["pred_err"] - perr) <= 1e-3 * perr
    assert abs(sel["pred_ms"] - pms) <= 1e-3 * pms


def test_out_of_range_features_fall_back_to_fixed_order(gpu):
    import torch
    from lsmps_vpm_v2_amd import FmmSolver, workloads as W
    f = W.random_clusters(30000, seed=6)
    dev = [torch.from_numpy(f[k]).cuda() for k in ("x", "y", "s", "omega", "active")]
    fixed = FmmSolver(p=12, max_par=48, max_src=48, precision="mixed", binom_mode=0,
                      list4_mode=2)
    u0, v0 = fixed.velocity_device(*dev)
    s = FmmSolver(p=12, max_par=48, max_src=48, precision="mixed", binom_mode=0, list4_mode=2)
    s.velocity_device(*dev)
    feats, _ = s.level_features()
    m = _model_for(feats.astype(np.float64), 7, 0.3)
    m["hi"][1] = m["lo"][1] + 1e-3  # cells-per-level range excludes this tree
    m["margin"] = 0.0
    _install(s, m, [48])
    s.set_opts(order_select=1)
    u1, v1 = s.velocity_device(*dev)
    sel = s.last_selection()
    assert sel["source"] == 2 and set(sel["p"][:len(feats)]) == {12}
    assert torch.equal(u0, u1) and torch.equal(v0, v1)


def test_trained_selector_gains_over_fixed_order(gpu):
    """The packaged model on fields it was not trained on: error <= 1e-4
    against the FP64 direct sum; at least 5 % faster than the same solver
    without the network (its fallback: fixed order 12, leaf rule 75/50) and
    than that order with the hand-tuned 48/48 rule; and within 8 % of the best
    fixed setting an exhaustive sweep finds (uniform p 7-9 x thresholds 48,
    64, 96 meeting 1e-4), which the network reaches without the sweep
    (measured: 1.00x on the flat plate, 1.04-1.06x on the shed-vortex field,
    where the model's threshold 96 is a little slower than the sweep's 64).
    Each setting's time is the median over 15 interleaved rounds of one solve
    (each call returns with its result complete) made after two untimed solves
    in a row, into preallocated outputs on a stream of its own, as a
    time-stepping code calls the solver."""
    import time
    import torch
    from lsmps_vpm_v2_amd import FmmSolver, selector as S, workloads as W
    if not os.path.exists(S.DEFAULT_MODEL):
        pytest.skip("no trained selector model in the package yet")
    for f in (W.shed_vortices(seed=99), W.boundary_layer(seed=17)):
        dev = [torch.from_numpy(f[k]).cuda() for k in ("x", "y", "s", "omega", "active")]
        q = torch.from_numpy(f["q"]).cuda()
        tidx = W.subsample(f, 20000, seed=3)
        td = torch.from_numpy(tidx).cuda()
        judge = FmmSolver()
        ex, ey = judge.direct_device(dev[0], dev[1], q, targets=td)
        ur, vr = ey.cpu().numpy(), -ex.cpu().numpy()

        def err_of(u, v):
            uu, vv = u.cpu().numpy()[tidx] - 1.0, v.cpu().numpy()[tidx]
            return float(np.sqrt(np.sum((uu - ur) ** 2 + (vv - vr) ** 2) /
                                 np.sum(ur ** 2 + vr ** 2)))

        # every solver is timed in the same rounds (one solve each per round,
        # in turn), so a slow stretch of the device hits them alike; the time
        # of each is its median over the rounds
        kw = dict(p=12, precision="mixed", binom_mode=0, list4_mode=2)
        solvers = {"fix": FmmSolver(**kw), "fix48": FmmSolver(max_par=48, max_src=48, **kw)}
        for p in (7, 8, 9):
            for t in (48, 64, 96):
                solvers[(p, t)] = FmmSolver(p=p, max_par=t, max_src=t, precision="mixed",
                                            binom_mode=0, list4_mode=2)
        learned = FmmSolver(order="learned", **kw)
        solvers["learned"] = learned
        outs = {}
        for k, s in solvers.items():
            for _ in range(2):
                outs[k] = s.velocity_device(*dev)
        torch.cuda.synchronize()
        ts = {k: [] for k in solvers}
        u_t, v_t = torch.empty_like(dev[0]), torch.empty_like(dev[0])
        with torch.cuda.stream(torch.cuda.Stream()):
            for _ in range(15):
                for k, s in solvers.items():
                    for _ in range(2):
                        s.velocity_device(*dev, u=u_t, v=v_t)
                    t0 = time.perf_counter()
                    s.velocity_device(*dev, u=u_t, v=v_t)
                    ts[k].append(time.perf_counter() - t0)
        med = {k: float(np.median(v)) for k, v in ts.items()}
        # the method without the network: the learned solver's own fallback,
        # fixed order 12 with the handle's leaf rule (the reference's 75/50)
        t_fix, e_fix = med["fix"], err_of(*outs["fix"])
        t_48 = med["fix48"]
        best = None
        for p in (7, 8, 9):
            for t in (48, 64, 96):
                if err_of(*outs[(p, t)]) <= 1e-4 and (best is None or med[(p, t)] < best[0]):
                    best = (med[(p, t)], p, t)
        t_lrn = med["learned"]
        e_lrn = err_of(*outs["learned"])
        sel = learned.last_selection()
        learned.set_opts(judge_every=1)  # one judged call: learned vs fixed p = 12
        learned.velocity_device(*dev)
        judged = learned.last_selection()
        print(f"{f['name']}: fixed p=12 (75/50) {t_fix * 1e3:.3f} ms err {e_fix:.2e} "
              f"[48/48: {t_48 * 1e3:.3f} ms]; best fixed p {best[1]} threshold {best[2]} "
              f"{best[0] * 1e3:.3f} ms; learned "
              f"{t_lrn * 1e3:.3f} ms err {e_lrn:.2e} (pred {sel['pred_err']:.2e}, judge "
              f"{judged['judge_err']:.2e}), threshold {sel['threshold']}, p {sel['p'][:12]}, "
              f"source {sel['source']}")
        assert sel["source"] == 1 and judged["source"] == 1
        assert e_lrn <= 1e-4 and judged["judge_err"] <= 0.9e-4
        assert t_lrn <= 0.95 * t_fix and t_lrn <= 0.95 * t_48
        assert t_lrn <= 1.08 * best[0]


def test_fixed_order_judge_rejects_a_bad_model(gpu):
    """A model that predicts negligible error everywhere picks the lowest
    order; the fixed-order judge (p = 12 on the same tree) sees the real
    difference, returns its own result and keeps the fixed order."""
    import torch
    from lsmps_vpm_v2_amd import FmmSolver, selector as S, workloads as W
    f = W.random_clusters(30000, seed=12)
    dev = [torch.from_numpy(f[k]).cuda() for k in ("x", "y", "s", "omega", "active")]
    kw = dict(p=12, max_par=48, max_src=48, precision="mixed", binom_mode=0, list4_mode=2)
    u0, v0 = FmmSolver(**kw).velocity_device(*dev)
    params = np.zeros(S.NP, np.float32)
    params[S.OFF["b3"][0]] = -60.0  # ln err^2 = -60 at every level and order
    big = np.full(9, 1e6, np.float32)
    model = S.Model(params=params, mu=np.zeros(9, np.float32), sd=np.ones(9, np.float32),
                    lo=-big, hi=big, p_min=4, p_max=16, thresholds=[48], target_err=1e-4)
    s = FmmSolver(order="learned", selector=model, **kw)  # default: judge fresh selections
    u1, v1 = s.velocity_device(*dev)
    sel = s.last_selection()
    print(f"judge: source {sel['source']} judge_err {sel['judge_err']:.2e}")
    assert sel["source"] == 3 and sel["judge_err"] > 1e-4
    assert torch.equal(u0, u1) and torch.equal(v0, v1)
    u2, v2 = s.velocity_device(*dev)  # stays on the fixed order
    assert s.last_selection()["source"] == 3
    assert torch.equal(u0, u2) and torch.equal(v0, v2)


def test_first_call_chooses_the_threshold_the_oracle_predicts(gpu):
    """A fresh learned solver re-selects the leaf threshold on its first call
    (every candidate tree built and scored on the device) and picks the one
    with the smallest predicted time under the oracle's evaluation of the same
    model on the same trees."""
    import torch
    from lsmps_vpm_v2_amd import FmmSolver, selector as S, workloads as W
    if not os.path.exists(S.DEFAULT_MODEL):
        pytest.skip("no trained selector model in the package yet")
    model = S.load_model()
    md = dict(params=model.params.astype(np.float64), mu=model.mu.astype(np.float64),
              sd=model.sd.astype(np.float64), lo=model.lo.astype(np.float64),
              hi=model.hi.astype(np.float64), margin=model.margin, p_min=model.p_min,
              p_max=model.p_max, target_err=model.target_err)
    f = W.boundary_layer()
    dev = [torch.from_numpy(f[k]).cuda() for k in ("x", "y", "s", "omega", "active")]
    pred = {}
    for t in model.thresholds:
        s = FmmSolver(p=12, max_par=t, max_src=t, precision="mixed", binom_mode=0, list4_mode=2)
        s.velocity_device(*dev)
        feats, _ = s.level_features()
        _, src, _, pms = M.select(md, feats.astype(np.float64), 12)
        if src == 1:
            pred[t] = pms
    assert pred, "the model should accept some candidate on this field"
    learned = FmmSolver(p=12, precision="mixed", binom_mode=0, list4_mode=2, order="learned")
    learned.set_opts(judge_every=-1)  # the threshold choice, not the judge
    learned.velocity_device(*dev)
    sel = learned.last_selection()
    best = min(pred, key=pred.get)
    print(f"oracle predictions {pred}; device chose {sel['threshold']} (source {sel['source']})")
    assert sel["reselected"] == 1 and sel["source"] == 1
    assert sel["threshold"] in pred
    assert pred[sel["threshold"]] <= pred[best] * 1.01
