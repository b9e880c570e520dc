#!/usr/bin/env python3
"""Benchmark: FMM particle-velocities/s on one MI355X.

Workload: the flat-plate boundary-layer field (779,280 particles, seeded,
lsmps_vpm_v2_amd.workloads.boundary_layer).  A step is one full velocity
solve of all particles with inputs already resident in HBM: binning, tree,
lists, every FMM stage and the write-back of (u, v) in caller order.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--path fp64|mixed]

Prints one JSON line.  ``value`` = particles / (seconds per step).  The
relative L2 error of the rotational velocity against an FP64 direct sum on
the GPU (seeded target sample, all sources) is reported next to it, as is
the plain FP64 path when the headline path is another one.  The CPU baseline
is the reference's own FMM (oracle/_ref, built from the reference sources)
timed on this host's cores on the same field.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FMM particle-velocities/sec at 780k particles, rel-L2 ≤1e-4 vs FP64 direct"
UNIT = "particle-velocities/s"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 (vector and matrix)
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 (vector and matrix)

# Solver options of each path (reference defaults for the plain FP64 path).
PATHS = {
    "fp64": dict(precision="fp64", p=10, max_par=75, max_src=50, binom_mode=1, list4_mode=1),
    # FP32 + MFMA M2L, exact binomials, list 4 by P2L, uniform order 8
    "mixed": dict(precision="mixed", p=8, max_par=48, max_src=48, binom_mode=0, list4_mode=2),
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def eval_flops(st, p):
    """Algorithmic flops of the evaluation kernel for one step: 11 per P2P
    pair, 8(p+1) per M2P (particle, list-3 cell), 8p per L2P particle."""
    return 11.0 * st["p2p_pairs"] + 8.0 * (p + 1) * st["m2p_evals"] + 8.0 * p * st["n"]


def time_path(torch, solver, dev, steps, warmup):
    for _ in range(warmup):
        solver.velocity_device(*dev)
    torch.cuda.synchronize()
    per_stage = []
    t0 = time.perf_counter()
    for _ in range(steps):
        solver.velocity_device(*dev)
        per_stage.append(solver.stats())
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    return dt, per_stage


def accuracy(torch, solver, field, dev_in, u, v, n_targets=20000):
    from lsmps_vpm_v2_amd import workloads
    tidx = workloads.subsample(field, n_targets)
    td = torch.from_numpy(tidx).to("cuda")
    q = torch.from_numpy(field["q"]).to("cuda")
    ex, ey = solver.direct_device(dev_in[0], dev_in[1], q, targets=td)
    u_ref, v_ref = ey.cpu().numpy(), -ex.cpu().numpy()
    uu = u.cpu().numpy()[tidx] - solver.opts.u_inf
    vv = v.cpu().numpy()[tidx] - solver.opts.v_inf
    num = np.sum((uu - u_ref) ** 2 + (vv - v_ref) ** 2)
    return float(np.sqrt(num / np.sum(u_ref ** 2 + v_ref ** 2))), tidx, (u_ref, v_ref)


def cpu_baseline(field, reps=3):
    """Reference FMM (oracle/_ref) on all host threads; falls back to the C
    restatement (oracle/liboracle_fmm2d.so, single thread) if the reference
    build is absent."""
    from oracle import oracle as O
    n = field["n"]
    cores = os.cpu_count() or 1
    try:
        cores = int(os.environ.get("OMP_NUM_THREADS", cores))
    except ValueError:
        pass
    times = []
    if O.ref_available():
        for _ in range(reps):
            t0 = time.perf_counter()
            O.ref_fmm_field(field["x"], field["y"], field["q"], field["active"], threads=cores)
            times.append(time.perf_counter() - t0)
        kind, used = "reference", cores
        sample = (f"reference TreeCell+fmm2D built from the reference sources, "
                  f"{n} particles (the full field), p=10, 75/50 leaf rule, median of {reps} solves")
    else:
        for _ in range(1):
            t0 = time.perf_counter()
            O.fmm_field(field["x"], field["y"], field["q"], field["active"], O.opts())
            times.append(time.perf_counter() - t0)
        kind, used = "port", 1
        sample = f"C restatement, {n} particles, single thread, 1 solve"
    t = statistics.median(times)
    return {"value": n / t, "unit": UNIT, "cores": used, "kind": kind, "sample": sample,
            "seconds_per_solve": t}


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--path", default="fp64", choices=sorted(PATHS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the multi-GPU "
              f"decomposition is a design (DESIGN.md), not a measured path.", file=sys.stderr)
        sys.exit(2)

    import torch
    from lsmps_vpm_v2_amd import FmmSolver, workloads

    field = workloads.boundary_layer()
    n = field["n"]
    log(f"[bench] workload {field['name']} n={n}")
    dev = [torch.from_numpy(field[k]).to("cuda") for k in ("x", "y", "s", "omega", "active")]
    torch.cuda.synchronize()

    opts = PATHS[args.path]
    solver = FmmSolver(timing=True, **opts)
    dt, stats = time_path(torch, solver, dev, args.steps, args.warmup)
    st = stats[-1]
    log(f"[bench] {args.path}: {dt * 1e3:.3f} ms/step, stages(ms) tree {st['ms_tree']:.3f} "
        f"lists {st['ms_lists']:.3f} up {st['ms_upward']:.3f} m2l {st['ms_m2l']:.3f} "
        f"down {st['ms_downward']:.3f} eval {st['ms_eval']:.3f}")

    u, v = solver.velocity_device(*dev)
    err = None
    if not args.no_accuracy:
        err, _, _ = accuracy(torch, solver, field, dev, u, v)
        log(f"[bench] rel-L2 vs FP64 direct (20k targets): {err:.3e}")

    # dominant kernel: the evaluation kernel (P2P + M2P + L2P), vector-ALU bound
    stage_ms = {k: float(np.mean([s[k] for s in stats])) for k in
                ("ms_tree", "ms_lists", "ms_upward", "ms_m2l", "ms_downward", "ms_eval")}
    eval_ms = stage_ms["ms_eval"]
    flops = eval_flops(st, opts["p"])
    peak = FP64_PEAK_TFLOPS if opts["precision"] == "fp64" else FP32_PEAK_TFLOPS
    achieved = flops / (eval_ms * 1e-3) / 1e12
    roofline = {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                "frac": achieved / peak, "traffic": None, "kernel": "k_eval64",
                "flops_per_launch": flops, "ms_per_launch": eval_ms}

    out = {
        "metric": METRIC, "value": n / dt, "unit": UNIT, "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if opts["precision"] == "fp64" else "f32",
        "data": "synthetic",
        "config": {"workload": field["name"], "n": n, "path": args.path, **opts,
                   "n_cells": st["n_cells"], "n_leaves": st["n_leaves"],
                   "leaf_levels": [st["min_leaf_level"], st["max_leaf_level"]]},
        "rel_l2_vs_fp64_direct": err,
        "stage_ms": stage_ms,
        "roofline": roofline,
    }
    if not args.no_cpu_baseline:
        log("[bench] timing the CPU baseline ...")
        out["cpu_baseline"] = cpu_baseline(field)
        log(f"[bench] cpu baseline {out['cpu_baseline']['seconds_per_solve']:.2f} s/solve")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
