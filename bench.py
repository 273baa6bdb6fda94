"""Throughput of the batched gas-phase kinetics path on one MI355X.

Workload (one "step"): CB05CL_AE5, 131072 independent cells (1820 columns x
72 levels), seed 1, T ~ U[220,320] K, half the cells at night (photolysis 0),
every cell advanced 3600 s with TR-BDF2 (dtmin = dt_in = 1e-20 s, dtmax 60 s,
atol 1e-12, rtol 1e-6; Newton atol 1e-18, rtol 1e-8, at most 20 iterations,
Jacobian refreshed every iteration).  Inputs are generated on the host and
copied to HBM before timing; a step restores the initial state (device copy)
and integrates every cell to t_end.

Forms:
  mixed (the metric): FP32 Jacobian assembly + sparse LU, FP64 state,
         residual, error estimate and step control; persistent work-queue
         launch, one kernel for the whole hour.
  fp64:  the same launch, all FP64 (measured in the same run unless
         --no-compare), with the max/median relative difference of the mixed
         result against it.
  fp64-plain: dense LU with partial pivoting (the oracle's arithmetic), one
         wavefront per cell, the reference's host interval loop (10 TR-BDF2
         iterations per launch); timed on --plain-cells cells.

Prints one JSON line on stdout; diagnostics go to stderr.  --copy-probe runs
only a 1 GiB device-to-device copy (a known-size kernel for calibrating
FETCH_SIZE / WRITE_SIZE counters).
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
KERNEL = "trbdf2_kernel"


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


def measured_traffic(workload, form):
    """HBM bytes per launch of the integrator from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary of the same workload and form (None when
    there is none)."""
    path = os.path.join(ROOT, "profiles", "trbdf2_kernel_cb05_131072_cells_hbm_traffic.json")
    if form != "mixed" or workload != "cb05_131072_cells_1h" or not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)["traffic_bytes_per_launch"]


def copy_probe():
    import torch
    n = 1 << 27  # 1 GiB of float64
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    log("copy probe: 1 GiB read + 1 GiB written")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--cells", type=int, default=131072)
    ap.add_argument("--form", choices=["mixed", "fp64", "fp64-plain"], default="mixed")
    ap.add_argument("--no-compare", action="store_true", help="skip the FP64 run of the same workload")
    ap.add_argument("--plain-cells", type=int, default=8192, help="cells timed in the fp64-plain form (0: skip)")
    ap.add_argument("--cpu-sample", type=int, default=16, help="cells in the CPU baseline sample (0: skip)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: min(16, os.cpu_count())")
    ap.add_argument("--copy-probe", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the path shards cells across GPUs with no exchange; "
              "multi-GPU runs are not part of this build", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe()
        return

    import torch
    T = importlib.import_module("tchem-atm_amd")
    S = importlib.import_module("tchem-atm_amd.scenarios")
    D = importlib.import_module("tchem-atm_amd.driver")
    import oracle

    kmd = T.KineticModelData(MECH)
    kmcd = kmd.createConstData()
    base = T.setScenarioConditions(MECH, kmd)
    m = kmd.nActive
    forms = {
        "mixed": dict(precision=T.PRECISION_MIXED, lu=T.LU_AUTO, persistent=True),
        "fp64": dict(precision=T.PRECISION_FP64, lu=T.LU_AUTO, persistent=True),
        "fp64-plain": dict(precision=T.PRECISION_FP64, lu=T.LU_DENSE, persistent=False),
    }

    def timed(form, n, steps, warmup):
        """Runs `steps` timed steps of the workload on n cells; returns
        (wall s/step, kernel-region s/step, launches/step, run)."""
        f = forms[form]
        w = S.config1(kmd, base[0], n=n, seed=1)
        run = D.DeviceRun(kmcd, w)

        def step(collect):
            run.reset()
            if f["persistent"]:
                return run.run_to_end(precision=f["precision"], schedule=T.SCHEDULE_PERSISTENT, lu=f["lu"],
                                      collect_stats=collect)
            return run.run_intervals(precision=f["precision"], schedule=T.SCHEDULE_GRID, lu=f["lu"],
                                     collect_stats=collect)

        for _ in range(warmup):
            step(False)
        torch.cuda.synchronize()
        stream = torch.cuda.current_stream()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        kernel_ms, launches = 0.0, 0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            run.reset()
            ev0.record(stream)
            launches += step(True)  # per-cell counters are written once per cell: no extra cost
            ev1.record(stream)
            ev1.synchronize()
            kernel_ms += ev0.elapsed_time(ev1)
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / steps
        return wall, kernel_ms / 1e3 / steps, launches / steps, run, w

    wall, kernel_s, launches, run, w = timed(args.form, args.cells, args.steps, args.warmup)
    value = w.cell_hours / wall
    st = run.total_stats
    t_out = run.t.cpu().numpy()
    if not np.all(t_out == w.tend):
        log("warning: not every cell reached t_end", int(np.sum(t_out != w.tend)))
    failed = int(st[:, 5].sum())
    acc = float(st[:, 0].sum())
    n_it = float(st[:, 3].sum()) / max(acc, 1.0)
    o = oracle.Oracle(MECH)
    f_rhs, f_jac = mechanism_flops(o)
    lu, rest = flops_per_step(m, n_it, f_rhs, f_jac)
    flop = acc * (lu + rest)
    mixed = forms[args.form]["precision"] == T.PRECISION_MIXED
    t_peak = acc * (lu / PEAK_FP32 + rest / PEAK_FP64) if mixed else flop / PEAK_FP64
    peak_eff = flop / t_peak
    achieved = flop / kernel_s
    log(f"form {args.form}: {wall:.2f} s wall per step, kernel region {kernel_s:.2f} s, {launches:.0f} launches; "
        f"steps/cell {acc / w.n:.0f}, Newton its/step {n_it:.2f}, failed cells {failed}")
    config = {"workload": "cb05_131072_cells_1h" if args.cells == 131072 else f"cb05_{args.cells}_cells_1h",
              "mechanism": "CB05CL_AE5 (74 active species, 187 reactions)", "cells": w.n, "seed": 1,
              "t_end_s": w.tend, "form": args.form, "accepted_steps_per_cell": acc / w.n,
              "newton_iterations_per_step": n_it, "failed_cells": failed}

    if args.form == "mixed" and not args.no_compare:
        wall64, _, _, run64, _ = timed("fp64", args.cells, 1, 0)
        v64 = w.cell_hours / wall64
        a = run.state.cpu().numpy()[:, 3:3 + m]
        b = run64.state.cpu().numpy()[:, 3:3 + m]
        err = np.abs(a - b) / np.maximum(np.abs(b), 1e-12)
        config.update({"fp64_value": v64, "mixed_over_fp64": value / v64,
                       "mixed_vs_fp64_relerr": {"max": float(err.max()), "median": float(np.median(err)),
                                                "q99": float(np.quantile(err, 0.99)), "floor": 1e-12}})
        log(f"fp64 same launch: {v64:.1f} cell-hours/s; mixed/fp64 {value / v64:.2f}; "
            f"rel diff max {err.max():.2e} median {np.median(err):.2e}")
    if args.form == "mixed" and args.plain_cells > 0:
        wallp, _, launchesp, runp, wp = timed("fp64-plain", args.plain_cells, 1, 0)
        vp = wp.cell_hours / wallp
        config.update({"fp64_plain_value": vp, "fp64_plain_cells": args.plain_cells,
                       "fp64_plain_launches": launchesp, "mixed_over_fp64_plain": value / vp})
        log(f"fp64-plain on {args.plain_cells} cells: {vp:.1f} cell-hours/s ({launchesp:.0f} launches)")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-hours/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32" if mixed else "f64",
        "data": "synthetic",
        "config": config,
        "roofline": {"bound": "mfma", "achieved": achieved / 1e12, "peak": peak_eff / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / peak_eff, "traffic": measured_traffic(config["workload"], args.form),
                     "flop_per_launch": flop / max(1.0, launches), "kernel": KERNEL,
                     "kernel_s_per_step": kernel_s},
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
