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

Other modes (one JSON line each, instead of the benchmark): --config3 times
the discrete adjoint against the forward solve, --config4 the uneven-batch
synthetic-mechanism workload, --precision-study the error study.

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


def config4_run(size_scale, out_path=None):
    """Uneven batches of the synthetic 250-species mechanism in one
    persistent launch (mixed form): cell-hours/s and per-batch completion
    times (last cell of the batch done, from the per-cell end timestamps)."""
    import tempfile

    import torch
    T = importlib.import_module("tchem-atm_amd")
    S = importlib.import_module("tchem-atm_amd.scenarios")
    D = importlib.import_module("tchem-atm_amd.driver")
    mech = S.synthetic_mechanism_yaml(os.path.join(tempfile.mkdtemp(), "synthetic250.yaml"))
    kmd = T.KineticModelData(mech)
    kmcd = kmd.createConstData()
    base = T.setScenarioConditions(mech, kmd)
    w, sizes = S.config4(kmd, base[0], size_scale=size_scale)
    run = D.DeviceRun(kmcd, w)
    run.reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run.run_to_end(precision=T.PRECISION_MIXED, schedule=T.SCHEDULE_PERSISTENT, collect_stats=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    st = run.total_stats
    start = run.t_start[run.t_start > 0].min()
    ends = np.split(run.t_end, np.cumsum(sizes)[:-1])
    done = np.array([(e.max() - start) / T.REALTIME_HZ for e in ends])
    res = {"workload": "synthetic250_48_batches", "size_scale": size_scale, "cells": int(sizes.sum()),
           "batch_sizes_min_max": [int(sizes.min()), int(sizes.max())], "cell_hours_per_s": w.cell_hours / wall,
           "wall_s": wall, "batch_completion_s_p50": float(np.percentile(done, 50)),
           "batch_completion_s_p99": float(np.percentile(done, 99)),
           "accepted_steps_per_cell": float(st[:, 0].mean()),
           "newton_iterations_per_step": float(st[:, 3].sum() / max(1, st[:, 0].sum())),
           "failed_cells": int(st[:, 5].sum())}
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
    ap.add_argument("--form", choices=["mixed", "fp64", "fp64-plain", "fp64-plain-graph"], default="mixed")
    ap.add_argument("--no-autotune", action="store_true",
                    help="persistent forms: default launch shape instead of the autotuner's")
    ap.add_argument("--no-compare", action="store_true", help="skip the FP64 run of the same workload")
    ap.add_argument("--plain-cells", type=int, default=8192, help="cells timed in the fp64-plain form (0: skip)")
    ap.add_argument("--cpu-sample", type=int, default=16, help="cells in the CPU baseline sample (0: skip)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: min(16, os.cpu_count())")
    ap.add_argument("--copy-probe", action="store_true")
    ap.add_argument("--precision-study", metavar="OUT", nargs="?", const="",
                    help="run the mixed/FP64-vs-oracle error study on config-0 inputs instead of the benchmark")
    ap.add_argument("--study-cells", type=int, default=1024)
    ap.add_argument("--config4", metavar="OUT", nargs="?", const="",
                    help="run the uneven-batch synthetic-mechanism workload instead of the benchmark")
    ap.add_argument("--config4-scale", type=float, default=1 / 64, help="batch-size scale of --config4")
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
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the path shards cells across GPUs with no exchange; "
              "multi-GPU runs are not part of this build", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe()
        return
    if args.config4 is not None:
        config4_run(args.config4_scale, args.config4 or None)
        return
    if args.config3 is not None:
        config3_run(args.config3_cells, args.config3 or None)
        return
    if args.e3sm is not None:
        e3sm_run(args.e3sm_tiles, args.e3sm or None)
        return
    if args.aerosol is not None:
        aerosol_run(args.aerosol_cells, args.aerosol_cpu_sample, args.cpu_threads or min(16, os.cpu_count() or 1),
                    args.aerosol or None, args.aerosol_particles)
        return
    if args.precision_study is not None:
        precision_study(args.study_cells, args.cpu_threads or min(16, os.cpu_count() or 1),
                        args.precision_study or None)
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
        "fp64-plain-graph": dict(precision=T.PRECISION_FP64, lu=T.LU_DENSE, persistent=False, graph=True),
    }
    tuned_shapes = {}

    def timed(form, n, steps, warmup):
        """Runs `steps` timed steps of the workload on n cells; returns
        (wall s/step, kernel-region s/step, launches/step, run)."""
        f = forms[form]
        w = S.config1(kmd, base[0], n=n, seed=1)
        run = D.DeviceRun(kmcd, w)
        shape = dict(lu=f["lu"], cells_per_block=0, blocks_per_cu=0)
        if f["persistent"] and not args.no_autotune:
            # launch shape from the autotuner on this batch (outside the timed region)
            run.reset(10_000_000)
            t0 = time.perf_counter()
            tuned = T.AtmosphericChemistry.autotune(run.tol_newton, run.tol_time, run.tadv, run.state, kmcd,
                                                    photo_rates=run.photo, precision=f["precision"],
                                                    schedule=T.SCHEDULE_PERSISTENT, lu=f["lu"])
            shape = dict(lu=tuned["lu"], cells_per_block=tuned["cells_per_block"],
                         blocks_per_cu=tuned["blocks_per_cu"])
            tuned_shapes[form] = dict(shape, tuning_s=time.perf_counter() - t0)
            log(f"autotuned {form}: {tuned_shapes[form]}")

        def step(collect):
            run.reset()
            if f["persistent"]:
                return run.run_to_end(precision=f["precision"], schedule=T.SCHEDULE_PERSISTENT,
                                      collect_stats=collect, **shape)
            if f.get("graph"):
                n_int = run.run_intervals_graph(precision=f["precision"], schedule=T.SCHEDULE_GRID, lu=f["lu"],
                                                intervals_per_launch=64)
                return (n_int + 63) // 64
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
        wallg, _, launchesg, rung, _ = timed("fp64-plain-graph", args.plain_cells, 1, 1)
        vg = wp.cell_hours / wallg
        same = bool(np.array_equal(runp.state.cpu().numpy(), rung.state.cpu().numpy()))
        config.update({"fp64_plain_value": vp, "fp64_plain_cells": args.plain_cells,
                       "fp64_plain_launches": launchesp, "fp64_plain_graph_value": vg,
                       "fp64_plain_graph_launches": launchesg, "fp64_plain_graph_identical": same,
                       "mixed_over_fp64_plain": value / max(vp, vg)})
        log(f"fp64-plain on {args.plain_cells} cells: {vp:.1f} cell-hours/s ({launchesp:.0f} interval launches); "
            f"captured loop {vg:.1f} cell-hours/s ({launchesg:.0f} graph launches, identical {same})")
    if tuned_shapes:
        config["autotuned"] = tuned_shapes

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
