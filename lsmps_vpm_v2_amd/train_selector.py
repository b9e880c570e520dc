"""Generate selector training data on the GPU and train the selection MLP
with the library's own kernels.

    python -m lsmps_vpm_v2_amd.train_selector [--orders 84] [--steps 6000]
        [--out lsmps_vpm_v2_amd/data/selector_mlp.bin] [--data sweeps.npz]
        [--from-data sweeps.npz]

Fields: ``selector.training_fields()`` (the uniform, boundary-layer, shed-
vortex and cluster fields at full size, plus 50k-particle subsamples); per
field, 5 leaf thresholds x ``orders`` order vectors.  One solve in ten is held
out to report how well the trained model predicts error and time.
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np


def log(msg):
    print(msg, flush=True)


def split(data, frac=0.1, seed=0):
    S = len(data["seg"]) - 1
    rng = np.random.Generator(np.random.PCG64(seed))
    hold = np.zeros(S, bool)
    hold[rng.choice(S, max(1, int(frac * S)), replace=False)] = True

    def take(mask):
        rows, seg = [], [0]
        for s in np.nonzero(mask)[0]:
            r = data["rows"][data["seg"][s]:data["seg"][s + 1]]
            rows.append(r)
            seg.append(seg[-1] + len(r))
        return dict(rows=np.concatenate(rows), seg=np.asarray(seg, np.int32),
                    y_err=data["y_err"][mask], y_time=data["y_time"][mask],
                    meta=[data["meta"][s] for s in np.nonzero(mask)[0]])
    return take(~hold), take(hold)


def predict(model, d):
    """Predicted ln err^2 and ln ms per solve with the float32 model on the GPU."""
    import torch
    from .solver import FmmSolver
    s = FmmSolver()
    xn = ((d["rows"] - model.mu) / model.sd).astype(np.float32)
    out = s.mlp_forward(torch.from_numpy(model.params).cuda(),
                        torch.from_numpy(xn).cuda()).cpu().numpy().astype(np.float64)
    pe, pt = [], []
    for a, b in zip(d["seg"][:-1], d["seg"][1:]):
        pe.append(np.log(np.exp(out[a:b, 0]).sum()))
        pt.append(np.log(np.exp(out[a:b, 1]).sum()))
    return np.asarray(pe), np.asarray(pt)


def main():
    from . import selector as S
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--orders", type=int, default=84)
    ap.add_argument("--steps", type=int, default=6000)
    ap.add_argument("--out", default=S.DEFAULT_MODEL)
    ap.add_argument("--data", default=None, help="write the sweep data here (.npz)")
    ap.add_argument("--from-data", default=None, help="train on saved sweep data")
    args = ap.parse_args()
    t0 = time.time()
    if args.from_data:
        z = np.load(args.from_data, allow_pickle=False)
        data = dict(rows=z["rows"], seg=z["seg"], y_err=z["y_err"], y_time=z["y_time"],
                    meta=[(str(a), int(b), int(c), [], float(d), float(e)) for a, b, c, d, e in
                          zip(z["m_name"], z["m_n"], z["m_t"], z["m_err"], z["m_ms"])])
    else:
        fields = S.training_fields()
        log(f"[train_selector] {len(fields)} fields: " + ", ".join(
            f"{f['name']}({f['n']})" for f in fields))
        data = S.sweep(fields, n_orders=args.orders, log=log)
        if args.data:
            os.makedirs(os.path.dirname(os.path.abspath(args.data)), exist_ok=True)
            np.savez_compressed(args.data, rows=data["rows"], seg=data["seg"],
                                y_err=data["y_err"], y_time=data["y_time"],
                                m_name=np.array([m[0] for m in data["meta"]]),
                                m_n=np.array([m[1] for m in data["meta"]]),
                                m_t=np.array([m[2] for m in data["meta"]]),
                                m_err=np.array([m[4] for m in data["meta"]]),
                                m_ms=np.array([m[5] for m in data["meta"]]))
    S_ = len(data["seg"]) - 1
    log(f"[train_selector] {S_} solves, {len(data['rows'])} level rows "
        f"({time.time() - t0:.1f} s)")
    tr, ho = split(data)
    model, hist = S.train(tr, steps=args.steps, log=log)
    pe, pt = predict(model, ho)
    ye, yt = ho["y_err"], ho["y_time"]
    err_fac = np.exp(0.5 * np.abs(pe - ye))   # factor between predicted and measured error
    time_fac = np.exp(np.abs(pt - yt))
    log(f"[train_selector] held-out {len(ye)} solves: error within x{np.median(err_fac):.2f} "
        f"(median), x{np.percentile(err_fac, 90):.2f} (90%); time within "
        f"x{np.median(time_fac):.3f} (median), x{np.percentile(time_fac, 90):.3f} (90%)")
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    model.save(args.out)
    log(f"[train_selector] wrote {args.out}; loss {hist[0]:.3f} -> {hist[-1]:.3f}; "
        f"total {time.time() - t0:.1f} s")
    return 0


if __name__ == "__main__":
    sys.exit(main())
