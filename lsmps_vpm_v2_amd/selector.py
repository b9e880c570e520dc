"""Learned choice of the expansion order per tree level and of the leaf
threshold (tree level) for the mixed-precision FMM.

Model (see include/vpmfmm.h, ``vpmfmm_selector``): per tree level a row of 9
features -- level, log2 cells, log2(particles / cell), log2 cell side,
log10(share of sum|q| under the level), log10(max|q| N / sum|q|), leaf
fraction, log2 leaf threshold, order p -- goes through 9 -> 64 tanh -> 64 tanh
-> 2.  The outputs (e, t) are the natural logs of the level's contribution to
the squared relative L2 velocity error and to the solve time in ms; a solve's
prediction is the sum of exp(.) over its levels.  The GPU chooses p per level
by minimising predicted time subject to predicted error <= target, and the
threshold by the smallest predicted time over the candidates; any level
feature outside the training range, or no feasible choice, falls back to the
fixed order.

This module holds the model file format, the sweep that generates training
data (FMM solves over orders and thresholds, errors against the FP64 direct
sum on the GPU), and training with the library's own MLP kernels.

Model file (little endian): b"VPMSEL01", int32 n_features, hidden, n_params,
p_min, p_max, n_thresholds, thresholds[8]; float32 target_err, margin,
mu[9], sd[9], lo[9], hi[9], params[n_params].
"""
from __future__ import annotations

import os
import struct
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib

F = _lib.SEL_FEATURES
H = _lib.SEL_HIDDEN
NP = _lib.SEL_PARAMS
MAGIC = b"VPMSEL01"
HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_MODEL = os.path.join(HERE, "data", "selector_mlp.bin")

# parameter blocks (offsets into the flat vector)
OFF = {"W1": (0, (F, H)), "b1": (F * H, (H,)), "W2": (F * H + H, (H, H)),
       "b2": (F * H + H + H * H, (H,)), "W3": (F * H + 2 * H + H * H, (H, 2)),
       "b3": (F * H + 4 * H + H * H, (2,))}
assert OFF["b3"][0] + 2 == NP


@dataclass
class Model:
    params: np.ndarray
    mu: np.ndarray
    sd: np.ndarray
    lo: np.ndarray
    hi: np.ndarray
    p_min: int = 4
    p_max: int = 16
    thresholds: list = field(default_factory=lambda: [32, 48, 64, 96, 128])
    target_err: float = 1e-4
    margin: float = 0.25

    def to_struct(self) -> _lib.Selector:
        s = _lib.Selector()
        s.n_features, s.hidden, s.n_params = F, H, NP
        s.p_min, s.p_max = self.p_min, self.p_max
        s.n_thresholds = len(self.thresholds)
        for k, t in enumerate(self.thresholds):
            s.thresholds[k] = int(t)
        s.target_err, s.margin = float(self.target_err), float(self.margin)
        for name in ("mu", "sd", "lo", "hi"):
            arr = getattr(s, name)
            for i, v in enumerate(np.asarray(getattr(self, name), np.float32)):
                arr[i] = float(v)
        return s

    def save(self, path: str) -> None:
        thr = list(self.thresholds) + [0] * (8 - len(self.thresholds))
        with open(path, "wb") as fh:
            fh.write(MAGIC)
            fh.write(struct.pack("<14i", F, H, NP, self.p_min, self.p_max, len(self.thresholds),
                                 *thr))
            fh.write(struct.pack("<2f", self.target_err, self.margin))
            for a in (self.mu, self.sd, self.lo, self.hi, self.params):
                fh.write(np.ascontiguousarray(a, "<f4").tobytes())

    @classmethod
    def load(cls, path: str) -> "Model":
        with open(path, "rb") as fh:
            raw = fh.read()
        if raw[:8] != MAGIC:
            raise ValueError(f"{path}: not a selector model file")
        ints = struct.unpack_from("<14i", raw, 8)
        nf, nh, npar, pmin, pmax, nt = ints[:6]
        if (nf, nh, npar) != (F, H, NP):
            raise ValueError(f"{path}: model shape {nf}/{nh}/{npar} != {F}/{H}/{NP}")
        thr = list(ints[6:6 + nt])
        off = 8 + 14 * 4
        target, margin = struct.unpack_from("<2f", raw, off)
        off += 8
        vec = np.frombuffer(raw, "<f4", count=4 * F + NP, offset=off).astype(np.float32)
        mu, sd, lo, hi = (vec[i * F:(i + 1) * F].copy() for i in range(4))
        return cls(params=vec[4 * F:].copy(), mu=mu, sd=sd, lo=lo, hi=hi, p_min=pmin,
                   p_max=pmax, thresholds=thr, target_err=float(target), margin=float(margin))


def load_model(path: str | None = None) -> Model:
    path = path or DEFAULT_MODEL
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path}: no trained selector model (see selector.train)")
    return Model.load(path)


def init_params(seed: int = 0, out_bias=(0.0, 0.0)) -> np.ndarray:
    """Glorot-uniform weights, zero biases except the output bias."""
    rng = np.random.Generator(np.random.PCG64(seed))
    p = np.zeros(NP, np.float32)
    for name, (o, shp) in OFF.items():
        if name.startswith("W"):
            lim = np.sqrt(6.0 / (shp[0] + shp[1]))
            p[o:o + shp[0] * shp[1]] = rng.uniform(-lim, lim, shp[0] * shp[1])
    p[OFF["b3"][0]:OFF["b3"][0] + 2] = out_bias
    return p


def rows_for_orders(feats: np.ndarray, orders) -> np.ndarray:
    """Feature rows of one solve: the level rows with the order column set
    (0 at levels 0 and 1, which carry no expansion)."""
    x = np.array(feats, np.float32, copy=True)
    for L in range(len(x)):
        x[L, F - 1] = 0.0 if L < 2 else float(orders[L])
    return x


# ---------------------------------------------------------------------------
# training data: FMM sweeps against the FP64 direct sum
# ---------------------------------------------------------------------------
def _field_on_gpu(fld):
    import torch
    dev = [torch.from_numpy(fld[k]).to("cuda") for k in ("x", "y", "s", "omega", "active")]
    q = torch.from_numpy(fld["q"]).to("cuda")
    return dev, q


def sweep(fields, thresholds=(32, 48, 64, 96, 128), n_orders=8, p_range=(4, 16), n_targets=50000,
          reps=3, seed=0, log=None):
    """FMM solves of every field at every threshold and ``n_orders`` order
    vectors per threshold (a base order with per-level jitter of +-2, plus
    uniform p_min and p_max); each solve's error is the relative L2 error of
    the rotational velocity against the FP64 direct sum at ``n_targets``
    seeded targets, its time the median of ``reps`` solves (device events).
    Returns dict(rows, seg, y_err, y_time, meta)."""
    import torch
    from . import workloads
    from .solver import FmmSolver
    rng = np.random.Generator(np.random.PCG64(seed))
    rows, seg, y_err, y_time, meta = [], [0], [], [], []
    judge = FmmSolver()
    for fi, fld in enumerate(fields):
        dev, q = _field_on_gpu(fld)
        tidx = workloads.subsample(fld, n_targets, seed=seed + fi)
        td = torch.from_numpy(tidx).to("cuda")
        ex, ey = judge.direct_device(dev[0], dev[1], q, targets=td)
        ur, vr = ey.cpu().numpy(), -ex.cpu().numpy()
        den = float(np.sum(ur ** 2 + vr ** 2))
        for t in thresholds:
            s = FmmSolver(p=p_range[1], max_par=t, max_src=t, precision="mixed", binom_mode=0,
                          list4_mode=2, timing=True)
            s.velocity_device(*dev)
            feats, _ = s.level_features()
            nlev = len(feats)
            vecs = [np.full(nlev, p_range[0]), np.full(nlev, p_range[1])]
            while len(vecs) < n_orders:
                base = rng.integers(p_range[0], p_range[1] + 1)
                v = np.clip(base + rng.integers(-2, 3, nlev), p_range[0], p_range[1])
                vecs.append(v)
            for v in vecs:
                s.set_level_orders(np.asarray(v, np.int32))
                times = []
                for _ in range(reps):
                    u, w = s.velocity_device(*dev)
                    times.append(s.stats()["ms_total"])
                uu = u.cpu().numpy()[tidx] - s.opts.u_inf
                ww = w.cpu().numpy()[tidx] - s.opts.v_inf
                err = float(np.sqrt(np.sum((uu - ur) ** 2 + (ww - vr) ** 2) / den))
                x = rows_for_orders(feats, v)
                rows.append(x)
                seg.append(seg[-1] + len(x))
                y_err.append(2.0 * np.log(max(err, 1e-12)))
                y_time.append(np.log(float(np.median(times))))
                meta.append((fld["name"], fld["n"], int(t), [int(a) for a in v], err,
                             float(np.median(times))))
            if log:
                log(f"[sweep] {fld['name']} n={fld['n']} t={t}: {len(vecs)} solves")
            s.close()
        del dev, q
    return dict(rows=np.concatenate(rows).astype(np.float32), seg=np.asarray(seg, np.int32),
                y_err=np.asarray(y_err, np.float32), y_time=np.asarray(y_time, np.float32),
                meta=meta)


def training_fields(seed: int = 0, n_sub: int = 10, sub_size: int = 50000):
    """Fields of the sweep: the uniform, boundary-layer, shed-vortex, cluster
    and wake (1M-particle von Karman street) fields, seeded ``sub_size``-
    particle random subsamples of them (sparser trees), and small fields
    (5k / 20k clusters, the Perlman vortex) for shallow trees."""
    from . import workloads as W
    full = [W.uniform_vortices(), W.boundary_layer(), W.shed_vortices(seed=11),
            W.random_clusters(200000, seed=3), W.von_karman(seed=6, n_total=1_000_000)]
    small = [W.random_clusters(5000, seed=9), W.random_clusters(20000, seed=10), W.perlman()]
    rng = np.random.Generator(np.random.PCG64(seed))
    subs = []
    for k in range(n_sub):
        src = full[k % len(full)]
        idx = np.sort(rng.choice(src["n"], min(sub_size, src["n"]), replace=False))
        d = {key: (val[idx] if isinstance(val, np.ndarray) and len(val) == src["n"] else val)
             for key, val in src.items()}
        d["n"] = len(idx)
        d["name"] = f"{src['name']}_sub{k}"
        subs.append(d)
    return full + subs + small


# ---------------------------------------------------------------------------
# training on the GPU with the library's kernels
# ---------------------------------------------------------------------------
def normalisation(rows: np.ndarray):
    mu = rows.mean(axis=0).astype(np.float32)
    sd = rows.std(axis=0).astype(np.float32)
    sd[sd < 1e-6] = 1.0
    return mu, sd


def train(data: dict, steps: int = 4000, lr: float = 3e-3, lr_min: float = 5e-5,
          w_time: float = 1.0, seed: int = 0, p_range=(4, 16), thresholds=None,
          target_err: float = 1e-4, margin: float = 0.25, solver=None, log=None) -> tuple:
    """Full-batch Adam on the GPU; returns (Model, loss history)."""
    import torch
    from .solver import FmmSolver
    rows = data["rows"]
    mu, sd = normalisation(rows)
    xn = ((rows - mu) / sd).astype(np.float32)
    nseg = len(data["seg"]) - 1
    y = np.stack([data["y_err"], data["y_time"]], axis=1).astype(np.float32).ravel()
    # output bias: the mean per-row share of each solve's total
    cnt = np.diff(data["seg"]).astype(np.float64)
    b_e = float(np.mean(data["y_err"] - np.log(cnt)))
    b_t = float(np.mean(data["y_time"] - np.log(cnt)))
    params = init_params(seed, (b_e, b_t))
    s = solver or FmmSolver()
    dev = "cuda"
    P = torch.from_numpy(params).to(dev)
    M = torch.zeros_like(P)
    V = torch.zeros_like(P)
    X = torch.from_numpy(xn).to(dev)
    S = torch.from_numpy(np.asarray(data["seg"], np.int32)).to(dev)
    Y = torch.from_numpy(y).to(dev)
    t0 = time.perf_counter()
    hist = s.mlp_train(P, M, V, X, S, Y, w_time, steps, 0, lr, lr_min)
    dt = time.perf_counter() - t0
    if log:
        log(f"[train] {steps} Adam steps over {len(rows)} rows / {nseg} solves in {dt:.2f} s; "
            f"loss {hist[0]:.4f} -> {hist[-1]:.4f}")
    feat = rows[:, :F - 1]
    lo = np.concatenate([feat.min(axis=0), [p_range[0]]]).astype(np.float32)
    hi = np.concatenate([feat.max(axis=0), [p_range[1]]]).astype(np.float32)
    thr = list(thresholds) if thresholds is not None else sorted({m[2] for m in data["meta"]})
    model = Model(params=P.cpu().numpy(), mu=mu, sd=sd, lo=lo, hi=hi, p_min=p_range[0],
                  p_max=p_range[1], thresholds=thr, target_err=target_err, margin=margin)
    return model, hist
