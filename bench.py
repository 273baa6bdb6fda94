"""Throughput of the batched gas-phase kinetics path on one MI355X.

Workload: CB05CL_AE5, 131072 independent cells (1820 columns x 72 levels),
seed 1, T ~ U[220,320] K, half the cells at night (photolysis 0), every cell
advanced one hour with TR-BDF2 (dtmin = dt_in = 1e-20 s, dtmax 60 s, atol
1e-12, rtol 1e-6; Newton atol 1e-18, rtol 1e-8, at most 20 iterations,
Jacobian refreshed every iteration).  Inputs are generated on the host and
copied to HBM before timing.

One step = one slice of the hour: --steps K cuts [0, 3600 s] into K equal
slices, and a step advances every cell from its own time to the slice end in
one persistent work-queue launch (the reference's host interval loop with the
interval as the slice: t_out / dt_out carried into the next slice's
TimeAdvance records, src/examples/TChem_AtmosphericChemistry.cpp:340-350).
The K timed steps together integrate the whole hour of all 131072 cells, so
value = 131072 cell-hours / (K x seconds per step).  --warmup W runs W
slices of a 4096-cell sample first.

Forms:
  mixed (the metric): FP32 Jacobian assembly + sparse LU, FP64 state,
         residual, error estimate and step control; persistent launch,
         autotuned launch shape.
  fp64:  the same launch, all FP64, measured on the same K slices, with the
         relative difference of the mixed result against it.
  fp64-plain: dense LU with partial pivoting (the oracle's arithmetic), one
         wavefront per cell, the reference's host interval loop (10 TR-BDF2
         iterations per launch), timed on a strided sample of the cells
         (--plain-cells), plus the same loop captured in a HIP graph.
  CPU:   the oracle on a strided sample of the cells, on the host's cores:
         the reference algorithm restated (forward-difference Jacobian, rate
         constants in every source-term call: cpu_baseline) and the
         analytic-Jacobian oracle (Newton iterations per step for the
         roofline, and the error of both GPU forms on the sample cells).
The side legs run under a wall-clock budget (--budget, seconds from start).

Other modes (one JSON line each, instead of the benchmark): --config3 times
the discrete adjoint against the forward solve, --config4 the uneven-batch
synthetic-mechanism workload, --precision-study the error study,
--cpu-study the CPU baseline on config 0 and a config-1 subset.

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


def host_info():
    """CPU model and the number of cores this process may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return model, usable


def default_threads():
    """Host threads for the CPU legs: the cores this process may use, at most
    16 (one GPU's share of the host)."""
    return max(1, min(16, host_info()[1]))


def cpu_leg(o, w, n_slices, mode, threads):
    """The oracle on the rows of workload w (the hour cut into n_slices
    intervals, as the GPU run): mode 1 = the reference algorithm restated
    (forward-difference Jacobian with m+1 source-term evaluations, all rate
    constants in every evaluation, a factorisation every Newton iteration,
    one cell per thread, static schedule, src/core/TChem_Util.hpp:71-73);
    mode 0 = analytic Jacobian.  Returns timing, final rows and counters."""
    import oracle
    t0 = time.perf_counter()
    st, t, dt, stats = oracle.run_sliced(o, w, w.state, w.photo, n_slices=n_slices, jacobian_mode=mode,
                                         nthreads=threads)
    el = time.perf_counter() - t0
    acc = stats["accepted"].astype(np.float64)
    nit = stats["newton_iterations"] / np.maximum(acc, 1.0)
    return dict(wall_s=el, value=w.cell_hours / el, state=st, t=t, stats=stats,
                steps_p50=float(np.percentile(acc, 50)), steps_p99=float(np.percentile(acc, 99)),
                newton_per_step=float(stats["newton_iterations"].sum() / max(1.0, acc.sum())),
                newton_per_step_p50=float(np.percentile(nit, 50)), newton_per_step_p99=float(np.percentile(nit, 99)),
                failed=int(stats["failed"].sum()))


def measured_traffic(workload, form):
    """HBM bytes per integrator launch from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary of this workload and form (None when
    there is none): rocprof counters cannot be read inside this process."""
    path = os.path.join(ROOT, "profiles", f"trbdf2_kernel_{workload}_{form}_hbm_traffic.json")
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        return json.load(f)["traffic_bytes_per_launch"], os.path.relpath(path, ROOT)


def rel_err(a, b, floor):
    return np.abs(a - b) / np.maximum(np.abs(b), floor)


def precision_study(cells, threads, out_path=None):
    """Mixed and FP64 GPU forms against the FP64 oracle on the config-0
    inputs (CB05CL_AE5, seed 0): per-species relative error
    |a - b| / max(|b|, floor) at t = 600, 1800 and 3600 s for rtol = 1e-4,
    1e-6 and 1e-8, with the number of failed (non-converged) cells."""
    import dataclasses

    import torch
    import oracle
    T = importlib.import_module("tchem-atm_amd")
    S = importlib.import_module("tchem-atm_amd.scenarios")
    D = importlib.import_module("tchem-atm_amd.driver")
    kmd = T.KineticModelData(MECH)
    kmcd = kmd.createConstData()
    base = T.setScenarioConditions(MECH, kmd)
    o = oracle.Oracle(MECH)
    m = kmd.nActive
    rows = []
    for rtol in (1e-4, 1e-6, 1e-8):
        w = dataclasses.replace(S.config0(kmd, base[0], n=cells, seed=0), rtol_time=rtol)
        runs = {}
        for name, prec in (("mixed", T.PRECISION_MIXED), ("fp64", T.PRECISION_FP64)):
            runs[name] = D.DeviceRun(kmcd, w)
            runs[name].reset()
        # oracle state carried across the three end times
        st = w.state.copy()
        t = np.full(w.n, w.tbeg)
        dt = np.full(w.n, w.dt_in)
        tn = np.array([w.atol_newton, w.rtol_newton])
        tt = np.tile([w.atol_time, w.rtol_time], (m, 1))
        for tend in (600.0, 1800.0, 3600.0):
            ta = np.zeros(w.n, oracle.TADV_DTYPE)
            ta["tbeg"], ta["tend"], ta["dt"], ta["dtmin"], ta["dtmax"] = t, tend, dt, w.dtmin, w.dtmax
            ta["max_num_newton_iterations"] = w.max_newton
            ta["num_time_iterations_per_interval"] = 10_000_000
            ta["jacobian_interval"] = w.jacobian_interval
            os_ = np.zeros(w.n, oracle.STATS_DTYPE)
            t0 = time.perf_counter()
            o.run_batch(tn, tt, ta, st, t, dt, photo=w.photo, stats=os_, nthreads=threads)
            el_o = time.perf_counter() - t0
            res = {}
            for name, prec in (("mixed", T.PRECISION_MIXED), ("fp64", T.PRECISION_FP64)):
                r = runs[name]
                # continue from each cell's own time and proposed step
                tbeg = r.t.cpu().numpy()
                dt0 = r.dt.cpu().numpy() if tend > 600.0 else w.dt_in
                tadv = T.time_advance_tensor(w.n, tbeg, tend, dt0, w.dtmin, w.dtmax, w.max_newton, 10_000_000,
                                             w.jacobian_interval, r.device)
                stats = T.stats_tensor(w.n, r.device)
                T.AtmosphericChemistry.runDeviceBatch(r.tol_newton, r.tol_time, None, tadv, r.state, r.t, r.dt,
                                                      r.state, kmcd, r.photo, stats, precision=prec,
                                                      schedule=T.SCHEDULE_PERSISTENT)
                torch.cuda.synchronize()
                g = r.state.cpu().numpy()[:, 3:3 + m]
                ok = (T.stats_from_tensor(stats)["failed"] == 0) & (os_["failed"] == 0)
                for floor in (1e-12, 1e-8):
                    e = (np.abs(g - st[:, 3:3 + m]) / np.maximum(np.abs(st[:, 3:3 + m]), floor))[ok]
                    res[(name, floor)] = (float(e.max()) if e.size else float("nan"), float(np.median(e)) if e.size else float("nan"))
                res[(name, "failed")] = int(np.sum(T.stats_from_tensor(stats)["failed"]))
            rows.append(dict(rtol=rtol, t=tend, oracle_failed=int(np.sum(os_["failed"])), oracle_s=el_o,
                             **{f"{n}_{k}": v for (n, k), v in res.items()}))
            log(f"rtol {rtol:g} t {tend:g}: " + ", ".join(f"{k} {v}" for k, v in rows[-1].items()
                                                          if k not in ("rtol", "t")))
    lines = ["precision study: CB05CL_AE5, config-0 inputs (seed 0), %d cells; per-species relative error" % cells,
             "|gpu - oracle| / max(|oracle|, floor) over cells that converged in both, FP64 oracle reference", "",
             f"{'rtol':>6} {'t[s]':>6} | {'mixed max':>10} {'mixed med':>10} {'max fl1e-8':>10} | "
             f"{'fp64 max':>10} {'fp64 med':>10} {'max fl1e-8':>10} | failed mixed/fp64/oracle"]
    for r in rows:
        lines.append(f"{r['rtol']:>6g} {r['t']:>6g} | {r['mixed_1e-12'][0]:>10.2e} {r['mixed_1e-12'][1]:>10.2e} "
                     f"{r['mixed_1e-08'][0]:>10.2e} | {r['fp64_1e-12'][0]:>10.2e} {r['fp64_1e-12'][1]:>10.2e} "
                     f"{r['fp64_1e-08'][0]:>10.2e} | {r['mixed_failed']}/{r['fp64_failed']}/{r['oracle_failed']}")
    text = "\n".join(lines)
    print(text, flush=True)
    if out_path:
        with open(out_path, "w") as f:
            f.write(text + "\n")


def config4_run(size_scale, out_path=None, forms=("mfma", "sparse"), coupling="uniform", slices=1):
    """Uneven batches of the synthetic 250-species mechanism (48 batches,
    sizes 10^U[log 256, log 65536] times size_scale) in one persistent
    launch of the mixed form, once per LU form: cell-hours/s, per-batch
    completion times (last cell of the batch done, from the per-cell end
    timestamps), and the forms' end states against each other.  With
    slices > 1 the hour runs as that many persistent launches (t and dt
    carried between them, a progress line each; slice ends log-spaced from
    0.1 s), for batch sets too large for one silent launch; batch completion
    times are then not reported."""
    import tempfile

    import torch
    T = importlib.import_module("tchem-atm_amd")
    S = importlib.import_module("tchem-atm_amd.scenarios")
    D = importlib.import_module("tchem-atm_amd.driver")
    mech = S.synthetic_mechanism_yaml(os.path.join(tempfile.mkdtemp(), "synthetic250.yaml"), coupling=coupling)
    kmd = T.KineticModelData(mech)
    kmcd = kmd.createConstData()
    base = T.setScenarioConditions(mech, kmd)
    w, sizes = S.config4(kmd, base[0], size_scale=size_scale)
    m = kmd.nActive
    lu = {"mfma": T.LU_MFMA, "sparse": T.LU_SPARSE, "dense": T.LU_DENSE}
    res = {"workload": "synthetic250_48_batches", "coupling": coupling, "size_scale": size_scale,
           "cells": int(sizes.sum()), "batch_sizes_min_max": [int(sizes.min()), int(sizes.max())], "forms": {}}
    log(f"config 4: {coupling} synthetic mechanism, {int(sizes.sum())} cells in 48 batches (scale {size_scale:g})")
    ends = {}
    run = D.DeviceRun(kmcd, w)
    for form in forms:
        if slices > 1:  # log-spaced slice ends: the stiff start holds most of the steps
            slice_ends = [float(x) for x in np.geomspace(0.1, w.tend, slices)]
            slice_ends[-1] = w.tend
            run.start_slices(slices, ends=slice_ends)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for k in range(slices):
                run.run_slice(k, precision=T.PRECISION_MIXED, schedule=T.SCHEDULE_PERSISTENT, lu=lu[form])
                torch.cuda.synchronize()
                log(f"  {form}: slice {k + 1}/{slices} done at {time.perf_counter() - t0:.1f} s")
            wall = time.perf_counter() - t0
            st = run.acc_stats.cpu().numpy()
            done = None
        else:
            run.reset()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            run.run_to_end(precision=T.PRECISION_MIXED, schedule=T.SCHEDULE_PERSISTENT, collect_stats=True,
                           lu=lu[form])
            torch.cuda.synchronize()
            wall = time.perf_counter() - t0
            st = run.total_stats
            start = run.t_start[run.t_start > 0].min()
            done = np.array([(e.max() - start) / T.REALTIME_HZ for e in np.split(run.t_end, np.cumsum(sizes)[:-1])])
        ends[form] = run.state.cpu().numpy()[:, 3:3 + m]
        r = {"cell_hours_per_s": w.cell_hours / wall, "wall_s": wall, "slices": slices,
             "accepted_steps_per_cell": float(st[:, 0].mean()),
             "newton_iterations_per_step": float(st[:, 3].sum() / max(1, st[:, 0].sum())),
             "failed_cells": int((st[:, 5] > 0).sum())}
        if done is not None:
            r["batch_completion_s_p50"] = float(np.percentile(done, 50))
            r["batch_completion_s_p99"] = float(np.percentile(done, 99))
        res["forms"][form] = r
        tail = (f", batch completion p50 {r['batch_completion_s_p50']:.1f} s p99 {r['batch_completion_s_p99']:.1f} s"
                if done is not None else "")
        log(f"  {form}: {r['cell_hours_per_s']:.2f} cell-hours/s ({wall:.1f} s){tail}, "
            f"steps/cell {r['accepted_steps_per_cell']:.0f}, failed {r['failed_cells']}")
    if len(ends) > 1:
        ref = ends[forms[-1]]
        for form in forms[:-1]:
            e8 = rel_err(ends[form], ref, 1e-8)
            res["forms"][form][f"max_rel_err_vs_{forms[-1]}_floor_1e-8"] = float(np.nanmax(e8))
    text = json.dumps(res)
    print(text, flush=True)
    if out_path:
        with open(out_path, "w") as f:
            f.write(text + "\n")


def config3_run(cells, out_path=None, max_steps=8192):
    """Cost of the discrete adjoint on the differentiation workload: CB05,
    config-0 draws with seed 2, advanced 1800 s.  Times the FP64 sparse
    persistent forward solve, the recording forward run and the backward
    sweep; loss L = sum_cells sum_k w_k log c_k(t_end), w ~ N(0, 1) over
    species >= 1e-12 ppm at t_end."""
    import torch
    T = importlib.import_module("tchem-atm_amd")
    S = importlib.import_module("tchem-atm_amd.scenarios")
    kmd = T.KineticModelData(MECH)
    kmcd = kmd.createConstData()
    base = T.setScenarioConditions(MECH, kmd)
    m, nspec, nR = kmd.nActive, kmd.nSpec, kmd.nReac
    tend = 1800.0
    w = S.config0(kmd, base[0], n=cells, seed=2)
    dev = "cuda"
    tn = torch.tensor([w.atol_newton, w.rtol_newton], dtype=torch.float64, device=dev)
    tt = torch.tensor(np.tile([w.atol_time, w.rtol_time], (m, 1)), device=dev)
    st0 = torch.tensor(w.state, device=dev)
    ph = torch.tensor(w.photo, device=dev)
    st = torch.empty_like(st0)
    t = torch.empty(cells, dtype=torch.float64, device=dev)
    dt = torch.empty_like(t)
    ws = T.Adjoint.workspace(kmcd, cells, max_steps)

    def reset():
        st.copy_(st0)
        t.fill_(0.0)
        dt.zero_()
        return T.time_advance_tensor(cells, 0.0, tend, w.dt_in, w.dtmin, w.dtmax, w.max_newton, 1 << 30, 1, dev)

    def plain():
        tadv = reset()
        T.AtmosphericChemistry.runDeviceBatch(tn, tt, None, tadv, st, t, dt, st, kmcd, photo_rates=ph,
                                              precision=T.PRECISION_FP64, schedule=T.SCHEDULE_PERSISTENT)

    def record():
        tadv = reset()
        T.Adjoint.forward(tn, tt, tadv, st, t, dt, ws, max_steps, kmcd, photo_rates=ph)

    def clock(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    plain()  # warm-up (module load)
    t_plain = clock(plain)
    log(f"config3: forward solve {t_plain:.3f} s")
    t_rec = clock(record)
    final = st.cpu().numpy()
    ns, _ = T.Adjoint.steps(ws, cells, max_steps, kmcd, with_h=False)
    ns = ns.cpu().numpy()
    c = final[:, 3:3 + m]
    wv = np.random.Generator(np.random.Philox(key=2)).standard_normal(m)
    wts = np.where(c >= 1e-12, wv, 0.0)
    lam = np.zeros((cells, nspec))
    lam[:, :m] = np.where(wts != 0.0, wts / np.where(wts != 0.0, c, 1.0), 0.0)
    lam_d = torch.tensor(lam, device=dev)
    lam_out = torch.empty_like(lam_d)
    grad = torch.empty((cells, nR), dtype=torch.float64, device=dev)
    t_back = clock(lambda: T.Adjoint.backward(lam_d, lam_out, grad, ws, max_steps, kmcd))
    g = grad.cpu().numpy()
    ok = ns >= 0
    gsum = g[ok].sum(0)
    res = {"workload": "cb05_config3_adjoint", "cells": cells, "t_end_s": tend, "max_steps": max_steps,
           "forward_solve_s": t_plain, "recording_forward_s": t_rec, "backward_s": t_back,
           "gradient_cost_vs_forward": (t_rec + t_back) / t_plain, "backward_vs_forward": t_back / t_plain,
           "accepted_steps_per_cell_mean": float(ns[ok].mean()), "accepted_steps_per_cell_max": int(ns[ok].max()),
           "cells_without_gradient": int((~ok).sum()), "loss_species_per_cell_mean": float((wts != 0).sum(1).mean()),
           "grad_norm_inf": float(np.abs(gsum).max()), "grad_finite": bool(np.isfinite(gsum).all()),
           "workspace_GiB": ws.numel() / 2 ** 30}
    text = json.dumps(res)
    print(text, flush=True)
    if out_path:
        with open(out_path, "w") as f:
            f.write(text + "\n")


def aerosol_run(cells, cpu_sample, threads, out_path=None, particles=1):
    """Single-particle SIMPOL.1 aerosol box model (AerosolChemistry): CB05CL_AE5
    with the SOA scheme, `particles` particles per cell (the file's 1 by
    default), 3600 s, one FP64 launch to t_end; cell-hours/s on the GPU and
    of the oracle on a sample of cells."""
    import torch
    T = importlib.import_module("tchem-atm_amd")
    S = importlib.import_module("tchem-atm_amd.scenarios")
    import oracle
    g = os.path.join(ROOT, "tests", "golden")
    chem, aero = os.path.join(g, "cb05_soa_gas.yaml"), os.path.join(g, "cb05_soa_aero.yaml")
    kmd = T.KineticModelData(chem)
    kmcd = kmd.createConstData()
    amd = T.AerosolModelData(aero, kmd)
    if particles != amd.nParticles:
        amd.setNumberofParticles(particles)
    amcd = amd.createConstData(kmcd)
    base = T.setScenarioConditions(chem, kmd)
    w, nc = S.aerosol_soa(kmd, base[0], amd.nSpec, amd.nParticles, n=cells)
    mt = amd.nUnknowns
    dev = "cuda"
    tn = torch.tensor([w.atol_newton, w.rtol_newton], dtype=torch.float64, device=dev)
    tt = torch.tensor(np.tile([w.atol_time, w.rtol_time], (mt, 1)), device=dev)
    st0 = torch.tensor(w.state, device=dev)
    ncd = torch.tensor(nc, device=dev)
    st = torch.empty_like(st0)
    t = torch.empty(cells, dtype=torch.float64, device=dev)
    dt = torch.empty_like(t)
    stats = T.stats_tensor(cells)

    def run():
        st.copy_(st0)
        t.fill_(0.0)
        dt.zero_()
        tadv = T.time_advance_tensor(cells, 0.0, w.tend, w.dt_in, w.dtmin, w.dtmax, w.max_newton, 1 << 30, 1, dev)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        T.AerosolChemistry.runDeviceBatch(tn, tt, None, tadv, st, ncd, t, dt, st, amcd, stats=stats)
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    log(f"aerosol: {cells} cells, {amd.nParticles} particle(s), {mt} unknowns per cell")
    log(f"aerosol: warm-up run {run():.2f} s")
    wall = run()
    log(f"aerosol: timed run {wall:.2f} s")
    s = T.stats_from_tensor(stats)
    res = {"metric": "cell_hours_per_s", "workload": f"cb05_soa_simpol_{particles}particle", "cells": cells,
           "unknowns_per_cell": mt, "t_end_s": w.tend, "value": w.cell_hours / wall, "wall_s": wall,
           "accepted_steps_per_cell": float(s["accepted"].mean()),
           "newton_iterations_per_step": float(s["newton_iterations"].sum() / max(1, s["accepted"].sum())),
           "failed_cells": int(s["failed"].sum()), "all_reached_t_end": bool((t.cpu().numpy() == w.tend).all())}
    def aerosol_oracle():
        o = oracle.AerosolOracle(chem, aero)
        if particles != o.a.nParticles:
            a = o.a
            a.nParticles = particles
            o = oracle.AerosolOracle(o.g, a)
        return o

    if cpu_sample > 0:
        o = aerosol_oracle()
        sample = np.ascontiguousarray(w.state[:cpu_sample])
        ta = np.zeros(cpu_sample, oracle.TADV_DTYPE)
        ta["tend"], ta["dt"], ta["dtmin"], ta["dtmax"] = w.tend, w.dt_in, w.dtmin, w.dtmax
        ta["max_num_newton_iterations"], ta["num_time_iterations_per_interval"] = w.max_newton, 1 << 30
        ta["jacobian_interval"] = 1
        to, dto = np.zeros(cpu_sample), np.zeros(cpu_sample)
        t0 = time.perf_counter()
        o.run_batch(np.array([w.atol_newton, w.rtol_newton]), np.tile([w.atol_time, w.rtol_time], (mt, 1)), ta,
                    sample, np.ascontiguousarray(nc[:cpu_sample]), to, dto, nthreads=threads)
        cpu_wall = time.perf_counter() - t0
        log(f"aerosol: oracle sample of {cpu_sample} cells {cpu_wall:.2f} s")
        res["cpu_baseline"] = {"value": cpu_sample * w.tend / 3600.0 / cpu_wall, "unit": "cell_hours_per_s",
                               "cores": threads, "kind": "port",
                               "sample": f"first {cpu_sample} cells of the workload, oracle, analytic Jacobian"}
        g_state = st.cpu().numpy()[:cpu_sample]
        cols = np.r_[3:3 + kmd.nActive, 3 + kmd.nSpec:amd.stateDim]
        err = np.abs(g_state[:, cols] - sample[:, cols]) / np.maximum(np.abs(sample[:, cols]), 1e-12)
        res["sample_max_rel_err_vs_oracle"] = float(err.max())
    bad = np.nonzero(s["failed"])[0]
    if len(bad):
        # do the cells the GPU gave up on fail in the oracle too?
        o = aerosol_oracle()
        ta = np.zeros(len(bad), oracle.TADV_DTYPE)
        ta["tend"], ta["dt"], ta["dtmin"], ta["dtmax"] = w.tend, w.dt_in, w.dtmin, w.dtmax
        ta["max_num_newton_iterations"], ta["num_time_iterations_per_interval"] = w.max_newton, 1 << 30
        ta["jacobian_interval"] = 1
        ost = np.ascontiguousarray(w.state[bad])
        ostats = np.zeros(len(bad), oracle.STATS_DTYPE)
        o.run_batch(np.array([w.atol_newton, w.rtol_newton]), np.tile([w.atol_time, w.rtol_time], (mt, 1)), ta,
                    ost, np.ascontiguousarray(nc[bad]), np.zeros(len(bad)), np.zeros(len(bad)), stats=ostats,
                    nthreads=threads)
        res["failed_cell_indices"] = [int(c) for c in bad]
        res["failed_cells_also_failed_by_oracle"] = int(ostats["failed"].sum())
    text = json.dumps(res)
    print(text, flush=True)
    if out_path:
        with open(out_path, "w") as f:
            f.write(text + "\n")


def e3sm_run(tiles, out_path=None):
    """E3SM box model (AtmosphericChemistryE3SM, UCI mechanism): the
    reference's 80-cell column scenario (per-cell photolysis rates and
    external forcing) tiled `tiles` times, 1800 s, one persistent launch to
    t_end per form (mixed and FP64 sparse).  There is no CPU restatement of
    the E3SM kinetics (its parity is pinned by the reference's goldens), so
    this mode has no CPU leg."""
    import torch
    T = importlib.import_module("tchem-atm_amd")
    g = os.path.join(ROOT, "tests", "golden")
    chem, scen = os.path.join(g, "uci_v2_test3.yaml"), os.path.join(g, "uci_col_input_conditions.yaml")
    kmd = T.KineticModelData(chem)
    kmcd = kmd.createConstData()
    base = T.setScenarioConditions(scen, kmd)
    n0 = base.shape[0]
    photo = T.setScenarioConditionsPhotolysisReactions(scen, kmd, n0)
    ext = T.setScenarioConditionsExternalForcing(scen, kmd, n0)
    cells = n0 * tiles
    m = kmd.nActive
    dev = "cuda"
    tn = torch.tensor([1e-18, 1e-8], dtype=torch.float64, device=dev)
    tt = torch.tensor(np.tile([1e-12, 1e-6], (m, 1)), device=dev)
    st0 = torch.tensor(np.tile(base, (tiles, 1)), device=dev)
    ph = torch.tensor(np.tile(photo, (tiles, 1)), device=dev)
    ex = torch.tensor(np.tile(ext, (tiles, 1)), device=dev)
    st = torch.empty_like(st0)
    t = torch.empty(cells, dtype=torch.float64, device=dev)
    dt = torch.empty_like(t)
    stats = T.stats_tensor(cells)
    res = {"metric": "cell_hours_per_s", "workload": "uci_e3sm_column_tiled", "cells": cells, "t_end_s": 1800.0,
           "mechanism": f"UCI v2 ({m} active species, {kmd.nReac} reactions)"}

    def run(prec):
        st.copy_(st0)
        t.fill_(0.0)
        dt.zero_()
        tadv = T.time_advance_tensor(cells, 0.0, 1800.0, 1e-20, 1e-20, 3600.0, 20, 1 << 30, 1, dev)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        T.AtmosphericChemistryE3SM.runDeviceBatch(tn, tt, None, tadv, st, ph, ex, t, dt, st, kmcd, stats=stats,
                                                  precision=prec, schedule=T.SCHEDULE_PERSISTENT)
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    for name, prec in (("mixed", T.PRECISION_MIXED), ("fp64", T.PRECISION_FP64)):
        run(prec)
        wall = run(prec)
        s = T.stats_from_tensor(stats)
        res[name] = {"value": cells * 1800.0 / 3600.0 / wall, "wall_s": wall,
                     "accepted_steps_per_cell": float(s["accepted"].mean()),
                     "failed_cells": int(s["failed"].sum()),
                     "all_reached_t_end": bool((t.cpu().numpy() == 1800.0).all())}
        if name == "fp64":
            fp64_state = st.cpu().numpy()
        else:
            mixed_state = st.cpu().numpy()
    c = slice(3, 3 + m)
    rel = np.abs(mixed_state[:, c] - fp64_state[:, c]) / np.maximum(np.abs(fp64_state[:, c]), 1e-12)
    res["mixed_vs_fp64_relerr"] = {"max": float(rel.max()), "median": float(np.median(rel))}
    res["value"] = res["mixed"]["value"]
    text = json.dumps(res)
    print(text, flush=True)
    if out_path:
        with open(out_path, "w") as f:
            f.write(text + "\n")


def cpu_study(threads, out_path=None, fd_cells=64, an_cells0=1024, an_cells1=2048):
    """The CPU baseline on the host's cores: the reference algorithm restated
    (forward-difference Jacobian) and the analytic oracle on config 0
    (CB05CL_AE5, seed 0, T ~ U[270,310] K) and on a strided subset of
    config 1 (seed 1, 131072 cells, T ~ U[220,320] K, half at night), each
    cell advanced 3600 s in one interval; cell-hours/s, accepted steps and
    Newton iterations per step (p50 / p99), failed cells."""
    import oracle
    T = importlib.import_module("tchem-atm_amd")
    S = importlib.import_module("tchem-atm_amd.scenarios")
    kmd = T.KineticModelData(MECH)
    base = T.setScenarioConditions(MECH, kmd)
    o = oracle.Oracle(MECH)
    model, usable = host_info()
    rows = []
    for name, make, n_total in (("config0", S.config0, 1024), ("config1", S.config1, 131072)):
        for mode, label, ncells in ((1, "reference algorithm (FD Jacobian)", fd_cells),
                                    (0, "analytic oracle", an_cells0 if name == "config0" else an_cells1)):
            idx = np.arange(0, n_total, max(1, n_total // ncells))[:ncells]
            w = make(kmd, base[0], seed=0 if name == "config0" else 1, cells=idx)
            r = cpu_leg(o, w, 1, mode, threads)
            rows.append(dict(workload=name, form=label, cells=int(idx.size), stride=int(max(1, n_total // ncells)),
                             value=r["value"], wall_s=r["wall_s"], steps_p50=r["steps_p50"], steps_p99=r["steps_p99"],
                             newton_per_step=r["newton_per_step"], newton_p50=r["newton_per_step_p50"],
                             newton_p99=r["newton_per_step_p99"], failed=r["failed"]))
            log(f"{name} {label}: {rows[-1]}")
    lines = [f"CPU baseline study: CB05CL_AE5, 3600 s per cell, one interval; host {model}, "
             f"{usable} usable cores, {threads} threads (OpenMP, static schedule, one cell per thread)", "",
             f"{'workload':8} {'form':34} {'cells':>6} {'cell-h/s':>9} {'wall s':>7} {'steps p50':>9} "
             f"{'p99':>7} {'Newton/step':>11} {'p50':>5} {'p99':>5} {'failed':>6}"]
    for r in rows:
        lines.append(f"{r['workload']:8} {r['form']:34} {r['cells']:>6} {r['value']:>9.3f} {r['wall_s']:>7.1f} "
                     f"{r['steps_p50']:>9.0f} {r['steps_p99']:>7.0f} {r['newton_per_step']:>11.3f} "
                     f"{r['newton_p50']:>5.2f} {r['newton_p99']:>5.2f} {r['failed']:>6}")
    text = "\n".join(lines)
    print(json.dumps({"cpu_study": rows, "cpu_model": model, "host_cores_usable": usable, "threads": threads}),
          flush=True)
    log(text)
    if out_path:
        with open(out_path, "w") as f:
            f.write(text + "\n")


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
    ap.add_argument("--steps", type=int, default=10, help="timed steps = slices of the simulated hour")
    ap.add_argument("--warmup", type=int, default=2, help="untimed slices on a 4096-cell sample first")
    ap.add_argument("--cells", type=int, default=131072)
    ap.add_argument("--form", choices=["mixed", "fp64"], default="mixed")
    ap.add_argument("--no-autotune", action="store_true",
                    help="persistent forms: default launch shape instead of the autotuner's")
    ap.add_argument("--no-compare", action="store_true", help="skip the FP64 run of the same workload")
    ap.add_argument("--kernel-only", action="store_true",
                    help="only the timed leg of --form (profiling runs): no FP64, plain or CPU legs")
    ap.add_argument("--plain-cells", type=int, default=16384, help="strided sample for the fp64-plain form (0: skip)")
    ap.add_argument("--graph-cells", type=int, default=4096, help="strided sample for the captured plain loop (0: skip)")
    ap.add_argument("--cpu-sample", type=int, default=0,
                    help="cells in the reference-algorithm CPU sample (0: one per host thread)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: the usable host cores, at most 16")
    ap.add_argument("--budget", type=float, default=420.0,
                    help="wall seconds from start after which optional legs are skipped")
    ap.add_argument("--copy-probe", action="store_true")
    ap.add_argument("--precision-study", metavar="OUT", nargs="?", const="",
                    help="run the mixed/FP64-vs-oracle error study on config-0 inputs instead of the benchmark")
    ap.add_argument("--study-cells", type=int, default=1024)
    ap.add_argument("--config4", metavar="OUT", nargs="?", const="",
                    help="run the uneven-batch synthetic-mechanism workload instead of the benchmark")
    ap.add_argument("--config4-scale", type=float, default=1 / 64, help="batch-size scale of --config4")
    ap.add_argument("--config4-forms", default="mfma,sparse", help="LU forms of --config4 (mfma, sparse, dense)")
    ap.add_argument("--config4-coupling", choices=["uniform", "cascade"], default="uniform")
    ap.add_argument("--config4-slices", type=int, default=1, help="run the hour of --config4 in this many launches")
    ap.add_argument("--config3", metavar="OUT", nargs="?", const="",
                    help="time the discrete adjoint (forward, recording forward, backward) instead of the benchmark")
    ap.add_argument("--config3-cells", type=int, default=4096)
    ap.add_argument("--aerosol", metavar="OUT", nargs="?", const="",
                    help="time the SIMPOL.1 aerosol box model instead (JSON line, optionally to OUT)")
    ap.add_argument("--aerosol-cells", type=int, default=16384)
    ap.add_argument("--aerosol-particles", type=int, default=1)
    ap.add_argument("--e3sm", metavar="OUT", nargs="?", const="",
                    help="time the E3SM box model on the tiled UCI column instead (JSON line, optionally to OUT)")
    ap.add_argument("--e3sm-tiles", type=int, default=8192)
    ap.add_argument("--aerosol-cpu-sample", type=int, default=128)
    ap.add_argument("--cpu-study", metavar="OUT", nargs="?", const="",
                    help="time the CPU baseline on config 0 and a config-1 subset instead of the benchmark")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the path shards cells across GPUs with no exchange; "
              "runs on more than one GPU are not part of this build", file=sys.stderr)
        sys.exit(2)
    threads = args.cpu_threads or default_threads()
    if args.copy_probe:
        copy_probe()
        return
    if args.cpu_study is not None:
        cpu_study(threads, args.cpu_study or None)
        return
    if args.config4 is not None:
        config4_run(args.config4_scale, args.config4 or None, tuple(args.config4_forms.split(",")),
                    args.config4_coupling, args.config4_slices)
        return
    if args.config3 is not None:
        config3_run(args.config3_cells, args.config3 or None)
        return
    if args.e3sm is not None:
        e3sm_run(args.e3sm_tiles, args.e3sm or None)
        return
    if args.aerosol is not None:
        aerosol_run(args.aerosol_cells, args.aerosol_cpu_sample, threads, args.aerosol or None,
                    args.aerosol_particles)
        return
    if args.precision_study is not None:
        precision_study(args.study_cells, threads, args.precision_study or None)
        return

    T0 = time.perf_counter()

    def left():
        return args.budget - (time.perf_counter() - T0)

    import torch
    T = importlib.import_module("tchem-atm_amd")
    S = importlib.import_module("tchem-atm_amd.scenarios")
    D = importlib.import_module("tchem-atm_amd.driver")

    kmd = T.KineticModelData(MECH)
    kmcd = kmd.createConstData()
    base = T.setScenarioConditions(MECH, kmd)
    m = kmd.nActive
    n, K = args.cells, max(1, args.steps)
    w = S.config1(kmd, base[0], n=n, seed=1)
    night = S.night_flags(n, 1, m, kmd.nPhoto)
    precisions = {"mixed": T.PRECISION_MIXED, "fp64": T.PRECISION_FP64}
    log(f"workload: {n} CB05 cells x 3600 s in {K} slices of {3600.0 / K:g} s; {int(night.sum())} night cells")

    def tune(run, precision):
        if args.no_autotune:
            return dict(lu=T.LU_AUTO, cells_per_block=0, blocks_per_cu=0, programs=T.PROGRAMS_AUTO)
        run.reset(10_000_000)
        t0 = time.perf_counter()
        r = T.AtmosphericChemistry.autotune(run.tol_newton, run.tol_time, run.tadv, run.state, kmcd,
                                            photo_rates=run.photo, precision=precision,
                                            schedule=T.SCHEDULE_PERSISTENT)
        shape = dict(lu=r["lu"], cells_per_block=r["cells_per_block"], blocks_per_cu=r["blocks_per_cu"],
                     programs=r["programs"])
        log(f"autotuned: {shape} ({time.perf_counter() - t0:.1f} s)")
        return shape

    def sliced(run, precision, shape, slices):
        """The hour in `slices` persistent launches; (wall s, [s per slice],
        [(longest cell time s, its index) per slice])."""
        run.start_slices(slices)
        stream = torch.cuda.current_stream()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(slices + 1)]
        longest = torch.zeros((slices, 2), dtype=torch.int64, device="cuda")
        st64 = run.stats.view(torch.int64).view(run.n, 5)  # [3] t_start, [4] t_end (device clock)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ev[0].record(stream)
        for k in range(slices):
            run.run_slice(k, precision=precision, schedule=T.SCHEDULE_PERSISTENT, **shape)
            el = st64[:, 4] - st64[:, 3]
            longest[k, 0] = el.max()
            longest[k, 1] = el.argmax()
            ev[k + 1].record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        lg = longest.cpu().numpy()
        return wall, [ev[k].elapsed_time(ev[k + 1]) / 1e3 for k in range(slices)], \
            [(float(x) / T.REALTIME_HZ, int(c)) for x, c in lg]

    # ---- headline form: warm-up on a sample, then K timed slices -----------------
    form = args.form
    prec = precisions[form]
    run = D.DeviceRun(kmcd, w)
    shape = tune(run, prec)
    if args.warmup > 0:
        wu = D.DeviceRun(kmcd, S.config1(kmd, base[0], seed=1, cells=np.arange(0, n, max(1, n // 4096))))
        wu.start_slices(K)
        for k in range(args.warmup):
            wu.run_slice(k % K, precision=prec, schedule=T.SCHEDULE_PERSISTENT, **shape)
        torch.cuda.synchronize()
        del wu
    wall, per_slice, longest_cells = sliced(run, prec, shape, K)
    value = w.cell_hours / wall
    kernel_s = float(sum(per_slice))
    acc = run.acc_stats.cpu().numpy()
    t_out = run.t.cpu().numpy()
    failed_mask = acc[:, 5] > 0
    steps_total = float(acc[:, 0].sum())
    n_it_gpu = float(acc[:, 3].sum()) / max(steps_total, 1.0)
    log(f"{form}: {value:.1f} cell-hours/s ({wall:.2f} s wall, kernel region {kernel_s:.2f} s, "
        f"slice max/mean {max(per_slice) / (kernel_s / K):.2f}); steps/cell {steps_total / n:.0f}, "
        f"Newton its/step {n_it_gpu:.3f}, failed cells {int(failed_mask.sum())}")
    if not np.all((t_out == w.tend) | failed_mask):
        log("warning: cells short of t_end:", int(np.sum((t_out != w.tend) & ~failed_mask)))
    state_main = run.state.cpu().numpy()
    config = {"workload": "cb05_131072_cells_1h" if n == 131072 else f"cb05_{n}_cells_1h",
              "mechanism": "CB05CL_AE5 (74 active species, 187 reactions)", "cells": n, "seed": 1,
              "night_cells": int(night.sum()), "t_end_s": w.tend, "form": form, "slices": K,
              "slice_s": w.tend / K, "launch_shape": shape, "accepted_steps_per_cell": steps_total / n,
              "accepted_steps_p50_p99": [float(np.percentile(acc[:, 0], 50)), float(np.percentile(acc[:, 0], 99))],
              "newton_iterations_per_step": n_it_gpu, "failed_cells": int(failed_mask.sum()),
              "failed_cell_indices": [int(i) for i in np.nonzero(failed_mask)[0][:64]],
              "slice_kernel_s": [round(x, 4) for x in per_slice],
              "slice_longest_cell_s": [round(x[0], 4) for x in longest_cells],
              "slice_longest_cell_index": [x[1] for x in longest_cells]}
    fidx = np.nonzero(failed_mask)[0]
    if fidx.size:
        log("failed cells (index: accepted, rejected, Newton failures):",
            ", ".join(f"{c}: {acc[c, 0]}/{acc[c, 1]}/{acc[c, 2]}" for c in fidx[:16]))
    del run
    torch.cuda.empty_cache()

    # ---- FP64 on the same launch and slices --------------------------------------
    if form == "mixed" and not args.no_compare and not args.kernel_only:
        run64 = D.DeviceRun(kmcd, w)
        shape64 = tune(run64, T.PRECISION_FP64)
        wall64, per64, _ = sliced(run64, T.PRECISION_FP64, shape64, K)
        v64 = w.cell_hours / wall64
        b = run64.state.cpu().numpy()[:, 3:3 + m]
        a = state_main[:, 3:3 + m]
        fail64 = run64.acc_stats[:, 5].cpu().numpy() > 0
        both = ~failed_mask & ~fail64
        e12, e8 = rel_err(a[both], b[both], 1e-12), rel_err(a[both], b[both], 1e-8)
        config.update({"fp64_value": v64, "fp64_launch_shape": shape64, "mixed_over_fp64": value / v64,
                       "fp64_failed_cells": int(fail64.sum()),
                       "failed_sets_equal": bool(np.array_equal(failed_mask, fail64)),
                       "mixed_vs_fp64_relerr": {
                           "floor_1e-12": {"max": float(e12.max()), "q99": float(np.quantile(e12, 0.99)),
                                           "median": float(np.median(e12))},
                           "floor_1e-8": {"max": float(e8.max()), "q99": float(np.quantile(e8, 0.99)),
                                          "median": float(np.median(e8))},
                           "cells_max_over_1e-4_floor_1e-12": int(np.sum(e12.max(axis=1) > 1e-4))}})
        log(f"fp64 same launch: {v64:.1f} cell-hours/s; mixed/fp64 {value / v64:.2f}; mixed vs fp64 max "
            f"{e12.max():.2e} (floor 1e-12) {e8.max():.2e} (floor 1e-8), median {np.median(e12):.1e}")
        state64 = run64.state.cpu().numpy()
        acc64 = run64.acc_stats.cpu().numpy()
        del run64
        torch.cuda.empty_cache()
    else:
        state64, fail64, acc64 = None, None, None

    # ---- plain FP64: dense LU, one wavefront per cell, host interval loop --------
    # The dense FP64 form computes in the oracle's arithmetic (bit for bit,
    # tests/test_gpu_parity.py), on the same K slices: its rows are the FP64
    # oracle's rows for the sampled cells, the reference for the error of the
    # mixed and sparse FP64 forms.
    if form == "mixed" and args.plain_cells > 0 and not args.kernel_only and left() > 150:
        stride = max(1, n // args.plain_cells)
        idx = np.arange(0, n, stride)[: args.plain_cells]
        wp = S.config1(kmd, base[0], seed=1, cells=idx)
        runp = D.DeviceRun(kmcd, wp)
        runp.start_slices(K)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        launches = 0
        for k in range(K):
            launches += runp.run_slice_intervals(k, iters=wp.iters_per_interval)
        torch.cuda.synchronize()
        wallp = time.perf_counter() - t0
        vp = wp.cell_hours / wallp
        oracle_rows = runp.state.cpu().numpy()[:, 3:3 + m]
        pacc = runp.acc_stats.cpu().numpy()
        ofail = pacc[:, 5] > 0
        config.update({"fp64_plain_value": vp, "fp64_plain_cells": int(idx.size),
                       "fp64_plain_sample": f"every {stride}th cell", "fp64_plain_launches": launches,
                       "mixed_over_fp64_plain": value / vp})
        log(f"fp64-plain on {idx.size} cells: {vp:.1f} cell-hours/s ({launches} interval launches, {wallp:.1f} s)")

        def against_oracle(rows, fail, stats):
            ok = ~ofail & ~fail[idx]
            a, b = rows[idx][ok][:, 3:3 + m], oracle_rows[ok]
            e12, e8 = rel_err(a, b, 1e-12), rel_err(a, b, 1e-8)
            tn = np.abs(a - b) / (w.atol_time + w.rtol_time * np.abs(b))
            # the 1e-4 bar of the metric, on every cell and on the cells whose
            # accept / reject sequence is the oracle's (elsewhere the step
            # sequences part and the cells differ within the integrator's
            # global error, as two FP64 forms do)
            same = (stats[idx, 0] == pacc[:, 0]) & (stats[idx, 1] == pacc[:, 1])
            e8s = e8.max(axis=1)[same[ok]]
            return {"cells": int(ok.sum()), "failed_sets_equal": bool(np.array_equal(fail[idx], ofail)),
                    "max_floor_1e-8": float(e8.max()), "max_floor_1e-12": float(e12.max()),
                    "median_floor_1e-12": float(np.median(e12)), "q99_floor_1e-12": float(np.quantile(e12, 0.99)),
                    "tolerance_normalised_max": float(tn.max()),
                    "cells_over_1e-9_floor_1e-12": int(np.sum(e12.max(axis=1) > 1e-9)),
                    "bar_1e-4_floor_1e-8": {
                        "met_on_all_cells": bool(e8.max() <= 1e-4),
                        "cells_over_bar": int(np.sum(e8.max(axis=1) > 1e-4)),
                        "cells_with_the_oracle_step_sequence": int(e8s.size),
                        "max_on_those": float(e8s.max()) if e8s.size else None,
                        "met_on_those": bool(e8s.size == 0 or e8s.max() <= 1e-4)}}

        acc_err = {form: against_oracle(state_main, failed_mask, acc)}
        if state64 is not None:
            acc_err["fp64"] = against_oracle(state64, fail64, acc64)
        config["error_vs_fp64_oracle_on_plain_sample"] = acc_err
        log("vs FP64 oracle on the plain sample: " + "; ".join(
            f"{k}: max {v['max_floor_1e-8']:.2e} (floor 1e-8) {v['max_floor_1e-12']:.2e} (floor 1e-12), "
            f"tolerance-normalised {v['tolerance_normalised_max']:.3f}; 1e-4 bar met on all cells "
            f"{v['bar_1e-4_floor_1e-8']['met_on_all_cells']} ({v['bar_1e-4_floor_1e-8']['cells_over_bar']} over), on the "
            f"{v['bar_1e-4_floor_1e-8']['cells_with_the_oracle_step_sequence']} cells with the oracle's step sequence "
            f"{v['bar_1e-4_floor_1e-8']['met_on_those']}" for k, v in acc_err.items()))
        del runp
        if args.graph_cells > 0 and left() > 90:
            gidx = np.arange(0, n, max(1, n // args.graph_cells))[: args.graph_cells]
            wg = S.config1(kmd, base[0], seed=1, cells=gidx)
            runs = []
            for graph in (False, True):
                r = D.DeviceRun(kmcd, wg)
                r.reset()
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                if graph:
                    nint = r.run_intervals_graph(precision=T.PRECISION_FP64, schedule=T.SCHEDULE_GRID,
                                                 lu=T.LU_DENSE, intervals_per_launch=64)
                else:
                    nint = r.run_intervals(precision=T.PRECISION_FP64, schedule=T.SCHEDULE_GRID, lu=T.LU_DENSE)
                torch.cuda.synchronize()
                runs.append((time.perf_counter() - t0, nint, r.state.cpu().numpy()))
            config.update({"fp64_plain_eager_vs_graph_cells": int(gidx.size),
                           "fp64_plain_eager_value": wg.cell_hours / runs[0][0],
                           "fp64_plain_graph_value": wg.cell_hours / runs[1][0],
                           "fp64_plain_graph_intervals": runs[1][1],
                           "fp64_plain_graph_identical": bool(np.array_equal(runs[0][2], runs[1][2]))})
            log(f"captured plain loop on {gidx.size} cells: eager {wg.cell_hours / runs[0][0]:.1f}, graph "
                f"{wg.cell_hours / runs[1][0]:.1f} cell-hours/s, identical {config['fp64_plain_graph_identical']}")

    # ---- CPU: the oracle on strided samples of the same cells ---------------------
    import oracle
    o = oracle.Oracle(MECH)
    f_rhs, f_jac = mechanism_flops(o)
    cpu = None
    n_it = n_it_gpu
    n_it_source = "GPU run (no CPU leg)"
    if not args.kernel_only and left() > 60:
        model, usable = host_info()
        ns = args.cpu_sample or threads
        idx = np.arange(0, n, max(1, n // ns))[:ns]
        ws = S.config1(kmd, base[0], seed=1, cells=idx)
        fd = cpu_leg(o, ws, K, 1, threads)
        log(f"cpu reference algorithm (FD Jacobian) on {ns} cells, {threads} threads: {fd['value']:.3f} "
            f"cell-hours/s ({fd['wall_s']:.1f} s)")
        na = 4 * threads
        aidx = np.arange(0, n, max(1, n // na))[:na]
        wa = S.config1(kmd, base[0], seed=1, cells=aidx)
        an = cpu_leg(o, wa, K, 0, threads)
        n_it, n_it_source = an["newton_per_step"], f"analytic oracle, {na}-cell strided sample"
        log(f"cpu analytic oracle on {na} cells: {an['value']:.3f} cell-hours/s; Newton its/step {n_it:.3f}")
        ok = (an["stats"]["failed"] == 0) & ~failed_mask[aidx]
        ea = rel_err(state_main[aidx][ok, 3:3 + m], an["state"][ok, 3:3 + m], 1e-12)
        sample_err = {f"{form}_max": float(ea.max()) if ea.size else None,
                      f"{form}_median": float(np.median(ea)) if ea.size else None}
        if state64 is not None:
            e64 = rel_err(state64[aidx][ok, 3:3 + m], an["state"][ok, 3:3 + m], 1e-12)
            sample_err.update({"fp64_max": float(e64.max()), "fp64_median": float(np.median(e64)),
                               "fp64_cells_over_1e-9": int(np.sum(e64.max(axis=1) > 1e-9))})
        cpu = {"value": fd["value"], "unit": "cell-hours/s", "cores": threads, "kind": "port",
               "sample": f"{ns} cells (every {max(1, n // ns)}th) of the same workload, 3600 s each in {K} "
                         f"slices, {fd['wall_s']:.1f} s wall; the reference algorithm restated in C "
                         f"(forward-difference Jacobian, rate constants in every source-term call)",
               "cpu_model": model, "host_cores_usable": usable,
               "accepted_steps_p50_p99": [fd["steps_p50"], fd["steps_p99"]],
               "newton_per_step_p50_p99": [fd["newton_per_step_p50"], fd["newton_per_step_p99"]],
               "failed_cells": fd["failed"],
               "analytic_oracle": {"value": an["value"], "cells": na, "wall_s": an["wall_s"],
                                   "accepted_steps_p50_p99": [an["steps_p50"], an["steps_p99"]],
                                   "newton_per_step": an["newton_per_step"], "failed_cells": an["failed"],
                                   "gpu_relerr_on_sample_floor_1e-12": sample_err}}

    # ---- roofline of the integrator launch ----------------------------------------
    lu_f, rest_f = flops_per_step(m, n_it, f_rhs, f_jac)
    flop = steps_total * (lu_f + rest_f)
    mixed = prec == T.PRECISION_MIXED
    t_peak = steps_total * (lu_f / PEAK_FP32 + rest_f / PEAK_FP64) if mixed else flop / PEAK_FP64
    peak = flop / t_peak
    achieved = flop / kernel_s
    traffic, traffic_src = measured_traffic(config["workload"], form)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-hours/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": wall / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32" if mixed else "f64",
        "data": "synthetic",
        "config": config,
        "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": KERNEL, "launches": K, "flop_per_launch": flop / K,
                     "avg_launch_s": kernel_s / K, "newton_iterations_per_step": n_it,
                     "newton_iterations_source": n_it_source,
                     "peak_note": ("precision-weighted vector peak: LU flops at the FP32 rate, the rest at FP64"
                                   if mixed else "FP64 vector peak")},
    }
    if cpu is not None:
        out["cpu_baseline"] = cpu
    bar = config.get("error_vs_fp64_oracle_on_plain_sample", {}).get(form, {}).get("bar_1e-4_floor_1e-8")
    if bar is not None:
        # the metric's accuracy bar (1e-4 relative, 1e-8 ppm floor) as measured on
        # the plain sample, stated beside the value rather than assumed
        out["accuracy_bar_1e-4"] = bar
    out["bench_wall_s"] = time.perf_counter() - T0
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
