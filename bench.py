#!/usr/bin/env python3
"""Benchmark: FMM particle-velocities/s on one MI355X.

Workload: the flat-plate boundary-layer field (779,280 particles, seeded,
lsmps_vpm_v2_amd.workloads.boundary_layer).  A step is one full velocity
solve of all particles with inputs already resident in HBM: binning, tree,
lists, order selection (learned path), every FMM stage and the write-back of
(u, v) in caller order.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--path learned|mixed|fp64]

Prints one JSON line.  ``value`` = particles / (seconds per step) of the
headline path (default: the learned-order FP32/MFMA path), timed without
stage events; the stage times and the evaluation kernel's duration come
from a second solver of the same options that records them.  The plain FP64
path (reference options: p = 10, 75/50 leaf rule) is timed beside it and the
ratio reported.  Errors: relative L2 of the rotational velocity against the
FP64 direct sum on the GPU (20,000 seeded targets, all sources) and against
the reference's own CPU FMM (all particles).  The CPU baseline is that
reference FMM (oracle/_ref, built from the reference sources) on this host's
cores.  ``inviscid_loop`` (not the metric) times the loop step of
vpmfmm_step_device on the same field: Euler advection, then the velocity at
the moved particles, a new tree every step.  ``fmm3d_path`` (not the
metric) times the 3-D treecode on a 778,873-particle vortex ring
(``--fmm3d-only`` prints that line alone).
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
FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 (vector; matrix is the same figure)
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 (vector; matrix is the same figure)

# Solver options of each path (the plain FP64 path uses the reference defaults).
PATHS = {
    "fp64": dict(precision="fp64", p=10, max_par=75, max_src=50, binom_mode=1, list4_mode=1),
    # FP32 + MFMA M2L, exact binomials, list 4 by P2L, uniform order 8
    "mixed": dict(precision="mixed", p=8, max_par=48, max_src=48, binom_mode=0, list4_mode=2),
    # the packaged MLP picks p per level and the leaf threshold; p = 12 fallback
    "learned": dict(precision="mixed", p=12, max_par=48, max_src=48, binom_mode=0, list4_mode=2,
                    order="learned"),
}
# the evaluation kernel of each path (template argument 0: the singular near field)
KERNEL = {"fp64": "k_eval64<0>", "mixed": "k_eval32<0>", "learned": "k_eval32<0>"}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def eval_flops(st):
    """Algorithmic flops of the evaluation kernel for one step: 11 per P2P
    (target, source) pair, 8 per M2P term (order + 1 per (particle, list-3
    cell)), 8 per L2P term (the leaf's order per particle)."""
    return 11.0 * st["p2p_pairs"] + 8.0 * st["m2p_terms"] + 8.0 * st["l2p_terms"]


def stage_work(solver, orders, fp64):
    """Algorithmic work of each stage on the solver's last tree, counted from
    its cells and lists (the inspection API) and the orders used per level:
    P2M 8(p+1) flops per active particle of a leaf; M2M 8(p+1)(p+2)/2 per
    active child link into a parent of level >= 1 (FP64) or >= 2 (mixed: its
    upward pass stops there); L2L the same per child link from a parent of
    level >= 2; M2L 8(p+1)^2 per list-2 translation (and per list-4 one on the
    FP64 path, whose M2L stage does list 4 by M2L); list build: each list
    entry written once (4 B) and each cell record read once (30 B)."""
    cells = solver.tree_cells()
    lev = cells["level"].astype(np.int64)
    leaf = cells["is_leaf"].astype(bool)
    act = cells["is_active"].astype(bool)
    p = np.asarray(orders, np.int64)
    pl = p[lev]
    n2 = np.zeros(len(lev), np.int64)
    n4 = np.zeros(len(lev), np.int64)
    for j in np.nonzero(lev >= 2)[0]:
        _, l2, _, l4 = solver.tree_lists(cells["id"][j], cap=256)
        n2[j], n4[j] = len(l2), len(l4)
    par = np.maximum(lev - 1, 0)
    pp = p[par]
    tri = 8.0 * (pp + 1) * (pp + 2) / 2.0
    up_floor = 1 if fp64 else 2
    p2m = 8.0 * float(np.sum(cells["src_num"][leaf & act] * (pl[leaf & act] + 1)))
    m2m = float(np.sum(tri[(lev >= 1) & (par >= up_floor) & act]))
    l2l = float(np.sum(tri[(lev >= 1) & (par >= 2)]))
    m2l = 8.0 * float(np.sum((n2 + (n4 if fp64 else 0)) * (pl + 1) ** 2))
    return {"p2m_flops": p2m, "m2m_flops": m2m, "l2l_flops": l2l, "m2l_flops": m2l,
            "list2_translations": int(n2.sum()), "list4_entries": int(n4.sum()),
            "n_cells": int(len(lev))}


def stage_rooflines(st, stage_ms, work, n, fp64):
    """Per-stage roofline entries and the time-weighted overall fraction
    sum_s t_s (achieved_s / peak_s) / sum_s t_s.  Bounds: the tree build and
    the list build by HBM (algorithmic bytes: 48 B per particle for the
    binning, read x, y, q and write them sorted; the list bytes of
    stage_work); M2L by the matrix cores (FP32 MFMA on the mixed path, FP64
    MFMA for list 2 on the FP64 path, whose list 4 runs on the vector ALU at
    the same FP64 peak); the other stages by the vector ALU (FP32 or FP64
    peak)."""
    vpeak = FP64_PEAK_TFLOPS if fp64 else FP32_PEAK_TFLOPS
    lists_bytes = 4.0 * (st["n_list1"] + st["n_list2"] + st["n_list3"] + st["n_list4"]) + \
        30.0 * st["n_cells"]
    spec = [
        ("tree", "ms_tree", "hbm", 48.0 * n, HBM_PEAK_GBS, "GB/s"),
        ("lists", "ms_lists", "hbm", lists_bytes, HBM_PEAK_GBS, "GB/s"),
        ("upward (P2M + M2M)", "ms_upward", "valu", work["p2m_flops"] + work["m2m_flops"],
         vpeak, "TFLOP/s"),
        ("M2L (list 2" + (" + list 4)" if fp64 else ")"), "ms_m2l", "mfma",
         work["m2l_flops"], vpeak, "TFLOP/s"),
        ("downward (L2L)", "ms_downward", "valu", work["l2l_flops"], vpeak, "TFLOP/s"),
        ("evaluation (P2P + M2P + L2P)", "ms_eval", "valu", eval_flops(st), vpeak, "TFLOP/s"),
    ]
    out, num, den = [], 0.0, 0.0
    for name, key, bound, work_s, peak, unit in spec:
        ms = stage_ms[key]
        scale = 1e9 if unit == "GB/s" else 1e12
        achieved = work_s / (ms * 1e-3) / scale if ms > 0 else 0.0
        frac = achieved / peak
        out.append({"stage": name, "bound": bound, "ms": ms, "work": work_s,
                    "work_unit": "bytes" if unit == "GB/s" else "flops", "achieved": achieved,
                    "peak": peak, "unit": unit, "frac": frac})
        num += ms * frac
        den += ms
    return out, (num / den if den > 0 else 0.0)


def time_path(torch, solver, dev, steps, warmup):
    """Mean seconds per step and the per-step statistics; the selection of
    the first call (which chooses the leaf threshold on the learned path) is
    kept on the solver as ``first_selection``.  The steps run on a stream of
    their own (the solver then runs on it directly) and write into output
    buffers allocated once, as a time-stepping code would."""
    u, v = torch.empty_like(dev[0]), torch.empty_like(dev[0])
    with torch.cuda.stream(torch.cuda.Stream(device=dev[0].device)):
        for k in range(warmup):
            solver.velocity_device(*dev, u=u, v=v)
            if k == 0:
                solver.first_selection = solver.last_selection()
        torch.cuda.synchronize()
        per_step = []
        t0 = time.perf_counter()
        for _ in range(steps):
            solver.velocity_device(*dev, u=u, v=v)
            per_step.append(solver.stats())
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
    return dt, per_step


def time_loop(torch, solver, dev, steps, warmup, dt=2e-3):
    """Mean milliseconds per step of the inviscid loop (vpmfmm_step_device:
    Euler advection with the current velocity, then the velocity at the moved
    particles) on copies of the positions, so each step builds a new tree."""
    x, y = dev[0].clone(), dev[1].clone()
    u, v = torch.empty_like(x), torch.empty_like(x)
    with torch.cuda.stream(torch.cuda.Stream(device=x.device)):
        solver.velocity_device(x, y, dev[2], dev[3], dev[4], u=u, v=v)
        for _ in range(warmup):
            solver.step_device(x, y, dev[2], dev[3], dev[4], u, v, dt)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            solver.step_device(x, y, dev[2], dev[3], dev[4], u, v, dt)
        torch.cuda.synchronize()
        sec = (time.perf_counter() - t0) / steps
    return {"ms_per_step": sec * 1e3, "steps": steps, "dt": dt, "scheme": "euler",
            "particles_moved_per_s": x.numel() / sec}


def time_host_path(FmmSolver, opts, field, steps, warmup):
    """The drop-in's own call: vpmfmm_velocity with host (pageable numpy)
    arrays, as VelocityCalc::get_velocity makes it -- the copies of x, y, s,
    omega, active in (33 B per particle) and u, v out (16 B) included."""
    s = FmmSolver(**opts)
    host = [np.ascontiguousarray(field[k]) for k in ("x", "y", "s", "omega", "active")]
    for _ in range(max(warmup, 1)):
        s.velocity(*host)
    t0 = time.perf_counter()
    for _ in range(steps):
        s.velocity(*host)
    sec = (time.perf_counter() - t0) / steps
    n = field["n"]
    return {"value": n / sec, "unit": UNIT, "ms_per_step": sec * 1e3, "steps": steps,
            "bytes_copied_per_step": 49 * n,
            "what": "vpmfmm_velocity (host pointers, pageable memory): the H2D copies of "
                    "x, y, s, omega, active, the solve, the D2H copies of u, v"}


def rel_l2(u, v, ur, vr):
    return float(np.sqrt(np.sum((u - ur) ** 2 + (v - vr) ** 2) / np.sum(ur ** 2 + vr ** 2)))


def direct_reference(torch, solver, field, dev, n_targets=20000):
    from lsmps_vpm_v2_amd import workloads
    tidx = workloads.subsample(field, n_targets)
    td = torch.from_numpy(tidx).to("cuda")
    q = torch.from_numpy(field["q"]).to("cuda")
    ex, ey = solver.direct_device(dev[0], dev[1], q, targets=td)
    return tidx, ey.cpu().numpy(), -ex.cpu().numpy()


def host_cores():
    """(threads used, physical cores from lscpu or None): the threads are the
    process's OpenMP share (OMP_NUM_THREADS, else the CPUs it may run on)."""
    import subprocess
    phys = None
    try:
        out = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True,
                             timeout=10).stdout
        pairs = {ln for ln in out.splitlines() if ln and not ln.startswith("#")}
        phys = len(pairs) or None
    except (OSError, subprocess.SubprocessError):
        pass
    threads = len(os.sched_getaffinity(0))
    try:
        threads = int(os.environ.get("OMP_NUM_THREADS", threads))
    except ValueError:
        pass
    return max(1, threads), phys


def cpu_baseline(field, reps=5, single_thread=True):
    """Reference FMM (oracle/_ref) on the host threads: one warm-up, then the
    median of `reps` solves of the full field; and one single-thread solve.
    Falls back to the C restatement (single thread) if the reference build is
    absent.  Returns (baseline dict, reference field (ex, ey) or None)."""
    from oracle import oracle as O
    n = field["n"]
    threads, phys = host_cores()
    times, ref = [], None
    if O.ref_available():
        O.ref_fmm_field(field["x"], field["y"], field["q"], field["active"], threads=threads)
        for _ in range(reps):
            t0 = time.perf_counter()
            ref = O.ref_fmm_field(field["x"], field["y"], field["q"], field["active"],
                                  threads=threads)
            times.append(time.perf_counter() - t0)
        kind, used = "reference", threads
        sample = (f"reference TreeCell::initializeTree + fmm2D::calcField built from the "
                  f"reference sources (g++ -O2 -fopenmp), the full {n}-particle field, p=10, "
                  f"75/50 leaf rule, {threads} OpenMP threads, median of {reps} solves after "
                  f"one warm-up")
    else:
        t0 = time.perf_counter()
        O.fmm_field(field["x"], field["y"], field["q"], field["active"], O.opts())
        times.append(time.perf_counter() - t0)
        kind, used = "port", 1
        sample = f"C restatement, {n} particles, single thread, 1 solve"
    t = statistics.median(times)
    out = {"value": n / t, "unit": UNIT, "cores": used, "kind": kind, "sample": sample,
           "seconds_per_solve": t, "solves_s": times, "host_physical_cores": phys}
    if kind == "reference" and single_thread:
        t0 = time.perf_counter()
        O.ref_fmm_field(field["x"], field["y"], field["q"], field["active"], threads=1)
        t1 = time.perf_counter() - t0
        O.set_threads(threads)  # the OpenMP setting is per calling thread: restore it
        out["single_thread"] = {"value": n / t1, "seconds_per_solve": t1, "cores": 1,
                                "sample": "the same solve on 1 thread, once"}
    return out, (None if ref is None else (ref[0], ref[1]))


def best_fixed(torch, FmmSolver, dev, tidx, ur, vr, u_inf, steps, warmup,
               orders=(6, 7, 8, 9, 10), thresholds=(32, 48, 64, 96, 128), target=1e-4):
    """The fixed-order mixed path swept over uniform orders x leaf thresholds
    (max_par = max_src): each setting's time per step (same timing as the
    headline) and rel-L2 against the FP64 direct sum at the same targets.
    Returns the fastest setting that meets `target` and the whole table."""
    table = []
    for t in thresholds:
        for p in orders:
            s = FmmSolver(precision="mixed", p=p, max_par=t, max_src=t, binom_mode=0,
                          list4_mode=2)
            dt, _ = time_path(torch, s, dev, steps, warmup)
            u, v = s.velocity_device(*dev)
            uu = u.cpu().numpy()[tidx] - u_inf
            vv = v.cpu().numpy()[tidx]
            table.append({"p": p, "threshold": t, "ms_per_step": dt * 1e3,
                          "rel_l2_vs_fp64_direct": rel_l2(uu, vv, ur, vr)})
    ok = [r for r in table if r["rel_l2_vs_fp64_direct"] <= target]
    best = min(ok, key=lambda r: r["ms_per_step"]) if ok else None
    return best, table


# Flops of the 3-D evaluation kernel as it computes them (an FMA counts 2):
# a direct pair 26 (offsets 3, |r|^2 5, 1/|r| 1, |r|^-3 2, cross product 9,
# scale and sum 6); a moment term 198 (1/|r| 1, powers 4, two scales 2,
# products 8, the 30 Taylor multipliers 60, the contraction 120, sum 3).
FMM3D_FLOPS_PAIR = 26.0
FMM3D_FLOPS_M2P = 198.0
FMM3D_METRIC = "3-D treecode particle-velocities/sec at 780k particles (vortex ring)"


def fmm3d_section(torch, steps, warmup, cpu, accuracy):
    """The 3-D treecode (the reference's VELOCITY_3D_TYPE 1,
    velocity_fmm_3d_fast) on the 778,873-particle vortex ring
    (workloads.vortex_ring) with the reference's DIM 3 options: time per
    solve with inputs in HBM, stage times and the evaluation kernel's
    roofline from a solver that records HIP events, lane use of the
    wavefront walk, the error against the 3-D direct sum (2,000 seeded
    targets, oracle), and the oracle on the host cores (the CPU baseline,
    whose velocities also check the GPU's to 1e-10)."""
    from lsmps_vpm_v2_amd import FmmSolver, workloads
    f = workloads.vortex_ring()
    n = f["n"]
    log(f"[bench] 3-D workload {f['name']} n={n}")
    dev = [torch.from_numpy(np.ascontiguousarray(f[k])).cuda()
           for k in ("x", "y", "z", "s", "wx", "wy", "wz")]
    act = torch.from_numpy(f["active"]).cuda()
    outs = [dev[0].new_empty(n) for _ in range(3)]
    s = FmmSolver()
    for _ in range(warmup):
        s.velocity3d_device(*dev, act, *outs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        s.velocity3d_device(*dev, act, *outs)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    st = s.fmm3d_stats()
    s_t = FmmSolver(timing=True)
    ms = {"ms_tree": [], "ms_moments": [], "ms_eval": []}
    for k in range(warmup + steps):
        s_t.velocity3d_device(*dev, act, *outs)
        if k >= warmup:
            stt = s_t.fmm3d_stats()
            for key in ms:
                ms[key].append(stt[key])
    stage_ms = {k: float(np.mean(v)) for k, v in ms.items()}
    flops = FMM3D_FLOPS_PAIR * st["pairs"] + FMM3D_FLOPS_M2P * st["m2p"]
    achieved = flops / (stage_ms["ms_eval"] * 1e-3) / 1e12
    u, v, w = (t.cpu().numpy() for t in outs)
    traffic, traffic_src = load_traffic("k3_eval", f["name"])
    log(f"[bench] 3-D: {dt * 1e3:.3f} ms/step; tree {stage_ms['ms_tree']:.3f} moments "
        f"{stage_ms['ms_moments']:.3f} eval {stage_ms['ms_eval']:.3f} ms; eval "
        f"{achieved:.2f} TFLOP/s ({achieved / FP64_PEAK_TFLOPS:.3f} of FP64)")
    out = {
        "metric": FMM3D_METRIC, "value": n / dt, "unit": UNIT, "ms_per_step": dt * 1e3,
        "steps": steps, "warmup": warmup, "dtype": "f64",
        "data": "synthetic (seeded vortex ring: Gaussian core, lattice inside the torus)",
        "config": {"workload": f["name"], "n": n, "crit_num": 30, "exp_tree": 1.0,
                   "support": 2.0, "n_cells": st["n_cells"], "n_leaves": st["n_leaves"],
                   "depth": st["depth"], "sorted_levels": st["sorted_levels"]},
        "stage_ms": stage_ms,
        "work": {"pairs": st["pairs"], "m2p": st["m2p"],
                 "lane_use_pairs": st["pairs"] / (64.0 * max(1, st["wave_pairs"])),
                 "lane_use_m2p": st["m2p"] / (64.0 * max(1, st["wave_m2p"]))},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS, "traffic": traffic,
                     "traffic_source": traffic_src,
                     "kernel": "k3_eval", "flops_per_launch": flops,
                     "ms_per_launch": stage_ms["ms_eval"],
                     "flops_note": f"{FMM3D_FLOPS_PAIR:.0f} per direct pair, "
                                   f"{FMM3D_FLOPS_M2P:.0f} per moment term",
                     "timing": "HIP events on the solver stream around the kernel, mean over "
                               "the steps of a second solver that records them"},
    }
    if accuracy or cpu:
        from oracle import oracle as O
        al = [f[k] * f["s"] ** 3 / (4 * np.pi) for k in ("wx", "wy", "wz")]
        rot = np.stack([u - 1.0, v, w], 1)
        if accuracy:
            rng = np.random.Generator(np.random.PCG64(7))
            tidx = np.sort(rng.choice(n, 2000, replace=False)).astype(np.int32)
            du, dv, dw = O.o3_direct(f["x"], f["y"], f["z"], *al, f["active"], targets=tidx)
            d = np.stack([du, dv, dw], 1)
            out["rel_l2_vs_fp64_direct"] = float(np.sqrt(((rot[tidx] - d) ** 2).sum() /
                                                         (d ** 2).sum()))
            log(f"[bench] 3-D rel-L2 vs the direct sum (2k targets): "
                f"{out['rel_l2_vs_fp64_direct']:.3e}")
        if cpu:
            threads, phys = host_cores()
            O.set_threads(threads)
            t0 = time.perf_counter()
            ou, ov, ow = O.o3_velocity(f["x"], f["y"], f["z"], f["s"], f["wx"], f["wy"],
                                       f["wz"], f["active"], O.o3_opts(), 1.0, 0.0, 0.0)
            tc = time.perf_counter() - t0
            orot = np.stack([ou - 1.0, ov, ow], 1)
            sc = np.sqrt(np.mean((orot ** 2).sum(1)))
            out["max_diff_vs_oracle"] = float(np.max(np.linalg.norm(rot - orot, axis=1)) / sc)
            out["cpu_baseline"] = {
                "value": n / tc, "unit": UNIT, "cores": threads, "kind": "port",
                "sample": f"the C restatement (oracle/fmm3d_oracle.c, OpenMP over targets) on "
                          f"the full {n}-particle ring, {threads} threads, one solve",
                "seconds_per_solve": tc, "host_physical_cores": phys}
            log(f"[bench] 3-D cpu baseline {tc:.3f} s/solve on {threads} threads; GPU vs "
                f"oracle max |d| / rms = {out['max_diff_vs_oracle']:.2e}")
    # the second 3-D path (VELOCITY_3D_TYPE 2, the TreeCell octree) on the same ring
    s2 = FmmSolver()
    for _ in range(warmup):
        s2.velocity3d_device(*dev, act, *outs, type=2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        s2.velocity3d_device(*dev, act, *outs, type=2)
    torch.cuda.synchronize()
    dt2 = (time.perf_counter() - t0) / steps
    st2 = s2.fmm3d_stats()
    t2 = {"ms_per_step": dt2 * 1e3, "value": n / dt2, "max_par": 50, "max_src": 30,
          "n_cells": st2["n_cells"], "n_leaves": st2["n_leaves"], "depth": st2["depth"],
          "pairs": st2["pairs"], "m2p": st2["m2p"]}
    log(f"[bench] 3-D type 2: {dt2 * 1e3:.3f} ms/step ({st2['n_cells']} cells)")
    if accuracy or cpu:
        from oracle import oracle as O
        u2, v2, w2 = (t.cpu().numpy() for t in outs)
        rot2 = np.stack([u2 - 1.0, v2, w2], 1)
        if accuracy:
            t2["rel_l2_vs_fp64_direct"] = float(np.sqrt(((rot2[tidx] - d) ** 2).sum() /
                                                        (d ** 2).sum()))
        if cpu:
            t0 = time.perf_counter()
            ou, ov, ow = O.o3_velocity(f["x"], f["y"], f["z"], f["s"], f["wx"], f["wy"],
                                       f["wz"], f["active"], O.o3_opts(type=2), 1.0, 0.0, 0.0)
            tc2 = time.perf_counter() - t0
            orot = np.stack([ou - 1.0, ov, ow], 1)
            sc = np.sqrt(np.mean((orot ** 2).sum(1)))
            t2["max_diff_vs_oracle"] = float(np.max(np.linalg.norm(rot2 - orot, axis=1)) / sc)
            t2["cpu_baseline_seconds_per_solve"] = tc2
    out["type2"] = t2
    return out


def load_traffic(kernel, workload):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary
    (FETCH_SIZE and WRITE_SIZE, scaled by a copy kernel of known size)."""
    path = os.path.join(ROOT, "profiles", f"{kernel.split('<')[0]}_{workload}_pmc.json")
    if not os.path.exists(path):
        return None, None
    with open(path) as fh:
        d = json.load(fh)
    return d.get("hbm_bytes_per_launch"), os.path.relpath(path, ROOT)


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--path", default="learned", choices=sorted(PATHS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison path")
    ap.add_argument("--no-loop", action="store_true",
                    help="skip the inviscid-loop timing (advection + velocity per step)")
    ap.add_argument("--no-host", action="store_true",
                    help="skip timing the host-pointer entry point (copies included)")
    ap.add_argument("--no-sweep", action="store_true",
                    help="skip the sweep of fixed orders x leaf thresholds (best fixed setting)")
    ap.add_argument("--no-fmm3d", action="store_true",
                    help="skip the 3-D treecode section (vortex ring)")
    ap.add_argument("--fmm3d-only", action="store_true",
                    help="time only the 3-D treecode and print its line")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB (the known-size "
                         "kernel that scales rocprofv3 FETCH_SIZE / WRITE_SIZE) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the multi-GPU "
              f"decomposition (Morton-range partition, top-tree all-gather, halo exchange) "
              f"is designed, not built or measured.", file=sys.stderr)
        sys.exit(2)

    import torch
    if args.copy_calibration:
        nb = args.copy_calibration << 20
        a = torch.empty(nb // 4, dtype=torch.float32, device="cuda").fill_(1.0)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": nb, "copy_bytes_written": nb, "copies": 5}))
        return
    from lsmps_vpm_v2_amd import FmmSolver, workloads

    if args.fmm3d_only:
        o3 = fmm3d_section(torch, args.steps, args.warmup, not args.no_cpu_baseline,
                           not args.no_accuracy)
        o3.update({"n_gpus": 1, "higher_is_better": True, "scaling": "weak",
                   "vs_baseline": None})
        print(json.dumps(o3), flush=True)
        return

    field = workloads.boundary_layer()
    n = field["n"]
    log(f"[bench] workload {field['name']} n={n}")
    dev = [torch.from_numpy(field[k]).to("cuda") for k in ("x", "y", "s", "omega", "active")]
    torch.cuda.synchronize()

    opts = PATHS[args.path]
    # the headline: the production configuration (no stage events)
    solver = FmmSolver(**opts)
    dt, stats = time_path(torch, solver, dev, args.steps, args.warmup)
    # stage times and the evaluation kernel's duration from a second solver
    # that records HIP events around the stages (they cost about 30 us per
    # solve, so they stay out of the headline)
    solver_t = FmmSolver(timing=True, **opts)
    dt_t, stats = time_path(torch, solver_t, dev, args.steps, args.warmup)
    st = stats[-1]
    sel = solver.last_selection()
    log(f"[bench] {args.path}: {dt * 1e3:.3f} ms/step ({dt_t * 1e3:.3f} with stage events), "
        f"stages(ms) tree {st['ms_tree']:.3f} "
        f"lists {st['ms_lists']:.3f} up {st['ms_upward']:.3f} m2l {st['ms_m2l']:.3f} "
        f"down {st['ms_downward']:.3f} eval {st['ms_eval']:.3f}")
    if args.path == "learned":
        log(f"[bench] selection: source {sel['source']} threshold {sel['threshold']} "
            f"p {sel['p'][:st['max_leaf_level'] + 1]} predicted err {sel['pred_err']:.2e} "
            f"ms {sel['pred_ms']:.3f}")
    fp64 = opts["precision"] == "fp64"
    sel_t = solver_t.last_selection()
    orders = sel_t["p"] if not fp64 else [opts["p"]] * 17
    work = stage_work(solver_t, orders, fp64)
    u, v = solver.velocity_device(*dev)
    u_h, v_h = u.cpu().numpy() - solver.opts.u_inf, v.cpu().numpy() - solver.opts.v_inf

    host = None
    if not args.no_host:
        host = time_host_path(FmmSolver, opts, field, max(3, args.steps // 4), 1)
        log(f"[bench] host-pointer path (vpmfmm_velocity, copies included): "
            f"{host['ms_per_step']:.3f} ms/step")

    loop = None
    if not args.no_loop:
        loop = time_loop(torch, solver, dev, args.steps, args.warmup)
        log(f"[bench] inviscid loop: {loop['ms_per_step']:.3f} ms/step "
            f"(advection + velocity, tree rebuilt every step)")

    out_fp64 = None
    if args.path != "fp64" and not args.no_fp64:
        s64 = FmmSolver(**PATHS["fp64"])
        dt64, _ = time_path(torch, s64, dev, args.steps, args.warmup)
        s64t = FmmSolver(timing=True, **PATHS["fp64"])
        _, st64 = time_path(torch, s64t, dev, args.steps, args.warmup)
        u64, v64 = s64.velocity_device(*dev)
        sms64 = {k: float(np.mean([s[k] for s in st64])) for k in
                 ("ms_tree", "ms_lists", "ms_upward", "ms_m2l", "ms_downward", "ms_eval")}
        w64 = stage_work(s64t, [PATHS["fp64"]["p"]] * 17, True)
        stg64, wt64 = stage_rooflines(st64[-1], sms64, w64, n, True)
        out_fp64 = {"value": n / dt64, "ms_per_step": dt64 * 1e3, "stage_ms": sms64,
                    "stage_roofline": stg64, "time_weighted_frac": wt64,
                    "config": PATHS["fp64"],
                    "_uv": (u64.cpu().numpy() - 1.0, v64.cpu().numpy())}
        log(f"[bench] fp64: {dt64 * 1e3:.3f} ms/step; speedup of {args.path} "
            f"{dt64 / dt:.2f}x")

    err = None
    sweep = None
    if not args.no_accuracy:
        tidx, ur, vr = direct_reference(torch, solver, field, dev)
        err = rel_l2(u_h[tidx], v_h[tidx], ur, vr)
        log(f"[bench] rel-L2 vs FP64 direct (20k targets): {err:.3e}")
        if out_fp64:
            uu, vv = out_fp64["_uv"]
            out_fp64["rel_l2_vs_fp64_direct"] = rel_l2(uu[tidx], vv[tidx], ur, vr)
        if args.path != "fp64" and not args.no_sweep:
            best, table = best_fixed(torch, FmmSolver, dev, tidx, ur, vr, solver.opts.u_inf,
                                     args.steps, args.warmup)
            sweep = {"best_fixed": best, "table": table,
                     "what": "fixed-order mixed path, uniform p in 6..10 x leaf threshold "
                             "(max_par = max_src) in {32, 48, 64, 96, 128}; the fastest "
                             "setting with rel-L2 <= 1e-4 against the FP64 direct sum"}
            if best:
                sweep["headline_speedup_vs_best_fixed"] = best["ms_per_step"] / (dt * 1e3)
                log(f"[bench] best fixed setting: p {best['p']} threshold {best['threshold']} "
                    f"{best['ms_per_step']:.3f} ms/step (rel-L2 "
                    f"{best['rel_l2_vs_fp64_direct']:.2e}); {args.path} is "
                    f"{sweep['headline_speedup_vs_best_fixed']:.3f}x that")

    stage_ms = {k: float(np.mean([s[k] for s in stats])) for k in
                ("ms_tree", "ms_lists", "ms_upward", "ms_m2l", "ms_downward", "ms_eval")}
    stage_ms["ms_per_step_with_events"] = dt_t * 1e3
    # dominant kernel: the evaluation kernel (P2P + M2P + L2P), vector-ALU bound
    eval_ms = stage_ms["ms_eval"]
    flops = eval_flops(st)
    peak = FP64_PEAK_TFLOPS if fp64 else FP32_PEAK_TFLOPS
    achieved = flops / (eval_ms * 1e-3) / 1e12
    traffic, traffic_src = load_traffic(KERNEL[args.path], field["name"])
    stages, weighted = stage_rooflines(st, stage_ms, work, n, fp64)
    roofline = {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                "frac": achieved / peak, "traffic": traffic, "kernel": KERNEL[args.path],
                "bound_note": "compute bound on the vector ALU, not on HBM; 'mfma' is the "
                              "compute-bound label of this field, and the peak is the MI355X "
                              + ("FP64" if fp64 else "FP32") + " vector figure",
                "compute_unit": "vector ALU (" + ("FP64" if fp64 else "FP32") + ")",
                "flops_per_launch": flops, "ms_per_launch": eval_ms,
                "timing": "HIP events on the solver stream around the kernel, mean over the "
                          "steps of a second solver that records stage events",
                "traffic_source": traffic_src,
                "stages": stages, "time_weighted_frac": weighted,
                "stage_work": work}
    log("[bench] stage rooflines: " + ", ".join(
        f"{e['stage'].split(' ')[0]} {e['achieved']:.2f} {e['unit']} ({e['frac']:.3f})"
        for e in stages) + f"; time-weighted {weighted:.3f}")

    cfg = {k: v for k, v in opts.items()}
    if args.path == "learned":
        # the options hold the fallback; the tree and orders actually used
        cfg = {k: v for k, v in cfg.items() if k not in ("p", "max_par", "max_src")}
        cfg["fallback"] = {"p": opts["p"], "max_par": opts["max_par"],
                           "max_src": opts["max_src"]}
        cfg["leaf_threshold"] = sel["threshold"]
        cfg["p_per_level"] = sel["p"][:st["max_leaf_level"] + 1]
    out = {
        "metric": METRIC, "value": n / dt, "unit": UNIT, "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if fp64 else "f32",
        "data": "synthetic (seeded flat-plate boundary-layer field)",
        "config": {"workload": field["name"], "n": n, "path": args.path, **cfg,
                   "n_cells": st["n_cells"], "n_leaves": st["n_leaves"],
                   "leaf_levels": [st["min_leaf_level"], st["max_leaf_level"]]},
        "rel_l2_vs_fp64_direct": err,
        "stage_ms": stage_ms,
        "roofline": roofline,
    }
    if args.path != "fp64":
        out["selection"] = {k: sel[k] for k in ("source", "threshold", "pred_err", "pred_ms")}
        out["selection"]["p_per_level"] = sel["p"][:st["max_leaf_level"] + 1]
        first = getattr(solver, "first_selection", None)
        if first and first["reselected"]:
            # once per run (and when n changes by > 10 %): not part of a step
            out["selection"]["threshold_choice_ms"] = first["ms_reselect"]
    if host:
        out["host_path"] = host
    if loop:
        out["inviscid_loop"] = loop
    if out_fp64:
        out["fp64_path"] = {k: val for k, val in out_fp64.items() if not k.startswith("_")}
        out["speedup_vs_fp64"] = out["value"] / out_fp64["value"]
    if sweep:
        out["fixed_order_sweep"] = sweep
    if not args.no_cpu_baseline:
        log("[bench] timing the CPU baseline ...")
        base, ref = cpu_baseline(field)
        out["cpu_baseline"] = base
        log(f"[bench] cpu baseline {base['seconds_per_solve']:.3f} s/solve on {base['cores']} "
            f"threads")
        if ref is not None:
            ur_all, vr_all = ref[1], -ref[0]
            out["rel_l2_vs_reference_fmm"] = rel_l2(u_h, v_h, ur_all, vr_all)
            if out_fp64:
                uu, vv = out_fp64["_uv"]
                out["fp64_path"]["rel_l2_vs_reference_fmm"] = rel_l2(uu, vv, ur_all, vr_all)
            log(f"[bench] rel-L2 vs reference FMM: {out['rel_l2_vs_reference_fmm']:.3e}")
    if not args.no_fmm3d:
        out["fmm3d_path"] = fmm3d_section(torch, args.steps, args.warmup,
                                          not args.no_cpu_baseline, not args.no_accuracy)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
