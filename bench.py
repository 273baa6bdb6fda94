"""Throughput of the batched gas-phase kinetics path on one MI355X.

Workload (one "step"): CB05CL_AE5, 131072 independent cells (1820 columns x
72 levels), seed 1, T ~ U[220,320] K, half the cells at night (photolysis 0),
every cell advanced 3600 s with TR-BDF2 (dtmin = dt_in = 1e-20 s, dtmax 60 s,
atol 1e-12, rtol 1e-6; Newton atol 1e-18, rtol 1e-8, at most 20 iterations,
Jacobian refreshed every iteration).  Inputs are generated on the host and
copied to HBM before timing; a step restores the initial state (device copy)
and integrates every cell to t_end.

Forms:
  mixed (default, the metric): FP32 iteration matrix + LU, FP64 state,
         residual, error estimate and step control; persistent work-queue
         launch, one kernel for the whole hour.
  fp64:  all-FP64, the reference's host interval loop (10 TR-BDF2 iterations
         per launch, relaunch until the mean time reaches 0.9999 t_end).

Prints one JSON line on stdout; diagnostics go to stderr.
"""
import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "simulated cell-hours/s, CB05 131072 cells: FP64 vs mixed/persistent at ≤1e-4 rel err"
MECH = os.path.join(ROOT, "tests", "golden", "cb05cl_ae5.yaml")
PEAK_FP64 = 78.6e12
PEAK_FP32 = 157.3e12


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def flops_per_step(m, n_it, f_rhs, f_jac):
    """Algorithmic FLOP of one accepted TR-BDF2 step of one cell: one LU of the
    m x m iteration matrix, one Jacobian, N_it Newton iterations (residual,
    two triangular solves, update) and the error estimate."""
    lu = 2.0 / 3.0 * m ** 3
    rest = m * m + f_jac + n_it * (2 * m * m + f_rhs + 4 * m) + f_rhs + 10 * m
    return lu, rest


def mechanism_flops(o):
    """F_rhs and F_jac of the mechanism from the oracle's tables: one multiply
    per reactant entry for the rates of progress plus one FMA per active
    stoichiometric entry; one derivative (two multiplies) per active reactant
    plus one FMA per Jacobian contribution."""
    k = o.k
    m = k.m
    half = k.maxSpecInReac // 2
    reactant_entries = int(np.sum(k.reacNreac))
    scatter = pairs = contrib = 0
    for r in range(k.nReac):
        act = [k.reacSidx[r, j] for j in range(k.reacNreac[r]) if k.reacSidx[r, j] < m]
        tgt = act + [k.reacSidx[r, half + j] for j in range(k.reacNprod[r]) if k.reacSidx[r, half + j] < m]
        scatter += len(tgt)
        pairs += len(act)
        contrib += len(act) * len(tgt)
    return reactant_entries + 2 * scatter, 2 * pairs + 2 * contrib


def cpu_baseline(w, sample, threads):
    """The reference algorithm restated on the CPU (forward-difference
    Jacobian with m+1 source-term evaluations, all rate constants in every
    evaluation, a fresh factorisation every Newton iteration, one cell per
    thread, static schedule), timed on the first `sample` cells."""
    import oracle
    o = oracle.Oracle(MECH)
    st = w.state[:sample].copy()
    ph = w.photo[:sample].copy()
    n = st.shape[0]
    t = np.full(n, w.tbeg)
    dt = np.full(n, w.dtmax)
    ta = np.zeros(n, oracle.TADV_DTYPE)
    ta["tbeg"], ta["tend"], ta["dt"], ta["dtmin"], ta["dtmax"] = w.tbeg, w.tend, w.dt_in, w.dtmin, w.dtmax
    ta["max_num_newton_iterations"] = w.max_newton
    ta["num_time_iterations_per_interval"] = w.iters_per_interval
    ta["jacobian_interval"] = w.jacobian_interval
    tn = np.array([w.atol_newton, w.rtol_newton])
    tt = np.tile([w.atol_time, w.rtol_time], (o.m, 1))
    t0 = time.perf_counter()
    for _ in range(1_000_000):
        o.run_batch(tn, tt, ta, st, t, dt, photo=ph, jacobian_mode=1, nthreads=threads)
        ta["tbeg"], ta["dt"] = t, dt
        if t.mean() > w.tend * 0.9999:
            break
    el = time.perf_counter() - t0
    return n * (w.tend - w.tbeg) / 3600.0 / el, el


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--cells", type=int, default=131072)
    ap.add_argument("--form", choices=["mixed", "fp64"], default="mixed")
    ap.add_argument("--cpu-sample", type=int, default=48, help="cells in the CPU baseline sample (0: skip)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: min(16, os.cpu_count())")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the path shards cells across GPUs with no exchange; "
              "multi-GPU runs are not part of this build", file=sys.stderr)
        sys.exit(2)

    import torch
    T = importlib.import_module("tchem-atm_amd")
    S = importlib.import_module("tchem-atm_amd.scenarios")
    D = importlib.import_module("tchem-atm_amd.driver")
    import oracle

    kmd = T.KineticModelData(MECH)
    kmcd = kmd.createConstData()
    base = T.setScenarioConditions(MECH, kmd)
    w = S.config1(kmd, base[0], n=args.cells, seed=1)
    run = D.DeviceRun(kmcd, w)
    torch.cuda.synchronize()
    m = kmd.nActive
    precision = T.PRECISION_MIXED if args.form == "mixed" else T.PRECISION_FP64

    def step(collect=False):
        run.reset()
        if args.form == "mixed":
            run.run_to_end(precision=precision, schedule=T.SCHEDULE_PERSISTENT, collect_stats=collect)
            return 1
        return run.run_intervals(precision=precision, schedule=T.SCHEDULE_GRID, collect_stats=collect)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    kernel_ms = 0.0
    launches = 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        run.reset()
        ev0.record(stream)
        if args.form == "mixed":
            # per-cell counters are written once per cell at its end: no extra cost
            run.run_to_end(precision=precision, schedule=T.SCHEDULE_PERSISTENT, collect_stats=True)
            launches += 1
        else:
            launches += run.run_intervals(precision=precision, schedule=T.SCHEDULE_GRID, collect_stats=True)
        ev1.record(stream)
        ev1.synchronize()
        kernel_ms += ev0.elapsed_time(ev1)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    value = w.cell_hours * args.steps / wall
    t_out = run.t.cpu().numpy()
    if not np.all(t_out == w.tend):
        log("warning: not every cell reached t_end", int(np.sum(t_out != w.tend)))

    st = run.total_stats  # counters of the last timed step
    failed = int(st[:, 5].sum())
    acc = float(st[:, 0].sum())
    n_it = float(st[:, 3].sum()) / max(acc, 1.0)
    o = oracle.Oracle(MECH)
    f_rhs, f_jac = mechanism_flops(o)
    lu, rest = flops_per_step(m, n_it, f_rhs, f_jac)
    flop = acc * (lu + rest)
    kernel_s = kernel_ms / 1e3 / args.steps  # region around the launches of one step
    if args.form == "mixed":
        t_peak = acc * (lu / PEAK_FP32 + rest / PEAK_FP64)
    else:
        t_peak = flop / PEAK_FP64
    peak_eff = flop / t_peak
    achieved = flop / kernel_s
    log(f"form {args.form}: {wall:.2f} s wall for {args.steps} step(s), kernel region {kernel_s:.2f} s/step, "
        f"{launches} launches; steps/cell {acc / w.n:.0f}, Newton its/step {n_it:.2f}, failed cells {failed}")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-hours/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32" if args.form == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": "cb05_config1", "cells": w.n, "seed": 1, "t_end_s": w.tend,
                   "form": "mixed precision, persistent work queue" if args.form == "mixed"
                   else "fp64, host interval loop",
                   "accepted_steps_per_cell": acc / w.n, "newton_iterations_per_step": n_it,
                   "failed_cells": failed},
        "roofline": {"bound": "mfma", "achieved": achieved / 1e12, "peak": peak_eff / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / peak_eff, "traffic": None,
                     "flop_per_launch": flop / max(1, launches // max(1, args.steps)),
                     "kernel": "trbdf2_kernel<float,2>" if args.form == "mixed" else "trbdf2_kernel<double,2>"},
    }
    if args.cpu_sample > 0:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        v, el = cpu_baseline(w, args.cpu_sample, threads)
        out["cpu_baseline"] = {"value": v, "unit": "cell-hours/s", "cores": threads, "kind": "port",
                               "sample": f"first {args.cpu_sample} cells of the same workload, 3600 s each, "
                                         f"{el:.1f} s wall"}
        log(f"cpu baseline: {v:.3f} cell-hours/s on {threads} threads ({el:.1f} s)")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
