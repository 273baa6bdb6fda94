"""HIP path against the CPU oracle — parity tests proper (need an MI355X).

All compute goes through the C ABI (libtchem_atm.so); the oracle is the
checker.  The plain FP64 form (dense LU with partial pivoting) computes in
the oracle's arithmetic — the same rate-constant functions, the reference's
serial summation order, the same fused operations and elimination order — so
it is compared bit for bit.  Other forms round differently; their error
metric per species is |gpu - oracle| / max(|oracle|, floor) with the floor
stated at each test, and the tolerance-normalised error
|gpu - oracle| / (atol + rtol |oracle|), which the integrator controls.
"""
import importlib

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu

FLOOR = 1e-12


def relerr(a, b, floor=FLOOR):
    return np.abs(a - b) / np.maximum(np.abs(b), floor)


@pytest.fixture(scope="module")
def cb05(gpu, pkg, cb05_oracle):
    kmd = pkg.KineticModelData(golden("cb05cl_ae5.yaml"))
    kmcd = kmd.createConstData()
    base = pkg.setScenarioConditions(golden("cb05cl_ae5.yaml"), kmd)
    return kmd, kmcd, base


@pytest.fixture(scope="module")
def scen():
    return importlib.import_module("tchem-atm_amd.scenarios")


@pytest.fixture(scope="module")
def drv():
    return importlib.import_module("tchem-atm_amd.driver")


def oracle_intervals(o, oracle_mod, w, stats=None, iters=None):
    n, m = w.n, o.m
    st = w.state.copy()
    t = np.full(n, w.tbeg)
    dt = np.full(n, w.dtmax)
    ta = np.zeros(n, oracle_mod.TADV_DTYPE)
    ta["tbeg"], ta["tend"], ta["dt"], ta["dtmin"], ta["dtmax"] = w.tbeg, w.tend, w.dt_in, w.dtmin, w.dtmax
    ta["max_num_newton_iterations"] = w.max_newton
    ta["num_time_iterations_per_interval"] = iters or w.iters_per_interval
    ta["jacobian_interval"] = w.jacobian_interval
    tn = np.array([w.atol_newton, w.rtol_newton])
    tt = np.tile([w.atol_time, w.rtol_time], (m, 1))
    total = np.zeros(n, oracle_mod.STATS_DTYPE)
    for _ in range(100000):
        s = np.zeros(n, oracle_mod.STATS_DTYPE)
        o.run_batch(tn, tt, ta, st, t, dt, photo=w.photo, stats=s)
        for f in ("accepted", "rejected", "newton_failures", "newton_iterations", "factorizations"):
            total[f] += s[f]
        total["failed"] |= s["failed"]
        ta["tbeg"], ta["dt"] = t, dt
        if t.mean() > w.tend * 0.9999:
            break
    return st, t, dt, total


def test_source_term_matches_oracle(cb05, cb05_oracle, scen, gpu, pkg):
    """Rate constants (portable exp / log / pow) and the serial source-term
    sum: the same bits as the oracle on 1024 config-0 cells."""
    kmd, kmcd, base = cb05
    torch = gpu
    w = scen.config0(kmd, base[0], n=1024, seed=0)
    st = torch.tensor(w.state, device="cuda")
    ph = torch.tensor(w.photo, device="cuda")
    om = torch.full((w.n, kmd.nActive), np.nan, dtype=torch.float64, device="cuda")
    pkg.SourceTerm.runDeviceBatch(st, ph, None, om, kmcd)
    ref = cb05_oracle.sourceterm_batch(w.state, w.photo)
    got = om.cpu().numpy()
    diff = got != ref
    print("source term entries differing from the oracle:", int(diff.sum()), "of", diff.size,
          "max rel", relerr(got, ref, 1e-300).max())
    np.testing.assert_array_equal(got, ref)
    # external sources are added; photo_rates=None uses the mechanism's A values
    ext = torch.ones_like(om)
    pkg.SourceTerm.runDeviceBatch(st, None, ext, om, kmcd)
    ref2 = cb05_oracle.sourceterm_batch(w.state, None) + 1.0
    assert (np.abs(om.cpu().numpy() - ref2) / (np.abs(ref2).max(axis=1, keepdims=True))).max() < 1e-14


def test_jacobian_matches_oracle(cb05, cb05_oracle, scen, gpu, pkg):
    kmd, kmcd, base = cb05
    torch = gpu
    w = scen.config0(kmd, base[0], n=256, seed=3)
    st = torch.tensor(w.state, device="cuda")
    ph = torch.tensor(w.photo, device="cuda")
    m = kmd.nActive
    jac = torch.full((w.n, m, m), np.nan, dtype=torch.float64, device="cuda")
    pkg.Jacobian.runDeviceBatch(st, ph, jac, kmcd)
    ref = cb05_oracle.jacobian_batch(w.state, w.photo)
    got = jac.cpu().numpy()
    print("Jacobian entries differing from the oracle:", int((got != ref).sum()), "of", got.size)
    np.testing.assert_array_equal(got, ref)


LU_FORMS = ("dense", "sparse")


def lu_code(pkg, name):
    return pkg.LU_DENSE if name == "dense" else pkg.LU_SPARSE


@pytest.mark.parametrize("lu", LU_FORMS)
def test_cb05cl_ae5_golden_on_gpu(cb05, gpu, pkg, lu):
    """The reference's ctest case run through runDeviceBatch on the GPU, both
    precisions, dense (partial pivoting) and sparse (static pivots) LU."""
    kmd, kmcd, base = cb05
    torch = gpu
    m = kmd.nActive
    gold = np.load(golden("cb05cl_ae5_full_gas.npy"))
    print("case: CB05CL_AE5 golden, lu", lu, flush=True)
    for precision in (pkg.PRECISION_FP64, pkg.PRECISION_MIXED):
        st = torch.tensor(base, device="cuda")
        tn = torch.tensor([1e-18, 1e-8], dtype=torch.float64, device="cuda")
        tt = torch.tensor(np.tile([1e-12, 1e-4], (m, 1)), device="cuda")
        tadv = pkg.time_advance_tensor(1, 0.0, 600.0, 1e-20, 1e-20, 10.0, 20, 10, 1)
        t = torch.zeros(1, dtype=torch.float64, device="cuda")
        dt = torch.full((1,), 10.0, dtype=torch.float64, device="cuda")
        rows = [np.concatenate([[-1, 0.0, 10.0], base[0]])]
        for it in range(2):
            pkg.AtmosphericChemistry.runDeviceBatch(tn, tt, None, tadv, st, t, dt, st, kmcd, precision=precision,
                                                    lu=lu_code(pkg, lu))
            pkg.AtmosphericChemistry.updateTimeAdvance(tadv, t, dt)
            torch.cuda.synchronize()
            rows.append(np.concatenate([[it, t.item(), dt.item()], st.cpu().numpy()[0]]))
        rows = np.array(rows)
        l2 = np.linalg.norm(gold, 2, 0)
        mx = np.where(l2 > 1e-23, l2, 1.0)
        crit = max((np.linalg.norm(rows - gold, o, 0) / mx).max() for o in (1, 2, np.inf))
        # the reference's ctest threshold for FP64; the FP32 factorisation leaves
        # Newton-tolerance-level (rtol_newton = 1e-8) differences
        assert crit < (1e-8 if precision == pkg.PRECISION_FP64 else 1e-6), (precision, crit)


@pytest.fixture(scope="module")
def oracle_config0_64(cb05, cb05_oracle, oracle_mod, scen):
    kmd, kmcd, base = cb05
    w = scen.config0(kmd, base[0], n=64, seed=0)
    return w, oracle_intervals(cb05_oracle, oracle_mod, w)


@pytest.fixture(scope="module", params=LU_FORMS)
def fp64_vs_oracle(request, cb05, oracle_config0_64, drv, gpu, pkg):
    kmd, kmcd, base = cb05
    w, (o_state, o_t, o_dt, o_stats) = oracle_config0_64
    lu = lu_code(pkg, request.param)
    print("case: config0 seed 0, 64 cells, fp64, lu", request.param, flush=True)
    run = drv.DeviceRun(kmcd, w)
    run.reset()
    run.run_intervals(precision=pkg.PRECISION_FP64, schedule=pkg.SCHEDULE_GRID, collect_stats=True, lu=lu)
    gpu.cuda.synchronize()
    g_state = run.state.cpu().numpy()
    g_t = run.t.cpu().numpy()
    g_stats = run.total_stats
    run.lu = lu
    return w, run, g_state, g_t, g_stats, o_state, o_t, o_stats


def test_fp64_integration_matches_oracle(fp64_vs_oracle, cb05, pkg):
    """config0 inputs (seed 0), 64 cells, 3600 s, the reference's interval
    loop, against the FP64 oracle.

    Dense form (the plain FP64 form): the same bits as the oracle in every
    cell — state, t_out, dt_out and every step / Newton counter — so the
    FP64 bar of 1e-9 relative per species holds with room to spare.

    Sparse form (static pivots in the fill-reducing order, other summation
    orders): rounds differently; the accept/reject decisions and Newton
    counts near their thresholds may then differ, which moves a cell within
    its integration tolerance.  Asserted: the tolerance-normalised error
    stays below 1 everywhere, and cells with the oracle's step sequence agree
    to 1e-4 relative above 1e-8 ppm."""
    w, run, g_state, g_t, g_stats, o_state, o_t, o_stats = fp64_vs_oracle
    m = cb05[0].nActive
    assert np.all(g_t == w.tend) and np.all(o_t == w.tend)
    assert g_stats[:, 5].sum() == 0 and o_stats["failed"].sum() == 0
    np.testing.assert_array_equal(g_state[:, 3 + m:], w.state[:, 3 + m:])  # constant species untouched
    np.testing.assert_array_equal(g_state[:, :3], w.state[:, :3])
    errs = relerr(g_state[:, 3:3 + m], o_state[:, 3:3 + m])
    err = errs.max(axis=1)
    same = (g_stats[:, 0] == o_stats["accepted"]) & (g_stats[:, 1] == o_stats["rejected"])
    print("lu", run.lu, "cells with the oracle's step sequence:", same.sum(), "of", w.n, "max rel err", err.max(),
          "bit-identical cells:", int(np.sum(np.all(g_state == o_state, axis=1))))
    if run.lu == pkg.LU_DENSE:
        for c in np.nonzero(np.any(g_state != o_state, axis=1))[0][:8]:
            print(f"  cell {c}: steps {g_stats[c, 0]}/{o_stats['accepted'][c]} rejected {g_stats[c, 1]}/"
                  f"{o_stats['rejected'][c]} newton {g_stats[c, 3]}/{o_stats['newton_iterations'][c]}")
        assert err.max() <= 1e-9
        np.testing.assert_array_equal(g_state, o_state)
        np.testing.assert_array_equal(g_t, o_t)
        np.testing.assert_array_equal(g_stats[:, 0], o_stats["accepted"])
        np.testing.assert_array_equal(g_stats[:, 1], o_stats["rejected"])
        np.testing.assert_array_equal(g_stats[:, 3], o_stats["newton_iterations"])
        np.testing.assert_array_equal(g_stats[:, 4], o_stats["factorizations"])
    else:
        tn = np.abs(g_state[:, 3:3 + m] - o_state[:, 3:3 + m]) / (w.atol_time + w.rtol_time * np.abs(o_state[:, 3:3 + m]))
        e8 = relerr(g_state[:, 3:3 + m], o_state[:, 3:3 + m], 1e-8).max(axis=1)
        print("  tolerance-normalised max", tn.max(), "rel err above 1e-8 ppm (same sequence):", e8[same].max())
        assert tn.max() < 1.0
        assert e8[same].max() < 1e-4


def test_fp64_dense_steps_match_oracle_step_by_step(cb05, cb05_oracle, oracle_mod, scen, gpu, pkg):
    """The dense FP64 form after 1, 2, 5, 20 and 100 TR-BDF2 iterations from
    config-1 states (cold and warm, day and night): the same bits as the
    oracle after every count (the first iteration count that differs names
    the step where the arithmetic parts)."""
    kmd, kmcd, base = cb05
    torch = gpu
    m = kmd.nActive
    w = scen.config1(kmd, base[0], seed=1, cells=np.arange(0, 131072, 4096))
    for iters in (1, 2, 5, 20, 100):
        print("case: config1 32 strided cells, dense fp64,", iters, "iterations", flush=True)
        st = torch.tensor(w.state, device="cuda")
        ph = torch.tensor(w.photo, device="cuda")
        tn = torch.tensor([w.atol_newton, w.rtol_newton], dtype=torch.float64, device="cuda")
        tt = torch.tensor(np.tile([w.atol_time, w.rtol_time], (m, 1)), device="cuda")
        tadv = pkg.time_advance_tensor(w.n, 0.0, w.tend, w.dt_in, w.dtmin, w.dtmax, w.max_newton, iters, 1)
        t = torch.zeros(w.n, dtype=torch.float64, device="cuda")
        dt = torch.zeros(w.n, dtype=torch.float64, device="cuda")
        pkg.AtmosphericChemistry.runDeviceBatch(tn, tt, None, tadv, st, t, dt, st, kmcd, photo_rates=ph,
                                                precision=pkg.PRECISION_FP64, lu=pkg.LU_DENSE)
        torch.cuda.synchronize()
        ost = w.state.copy()
        ot, odt = np.zeros(w.n), np.zeros(w.n)
        ta = np.zeros(w.n, oracle_mod.TADV_DTYPE)
        ta["tend"], ta["dt"], ta["dtmin"], ta["dtmax"] = w.tend, w.dt_in, w.dtmin, w.dtmax
        ta["max_num_newton_iterations"], ta["num_time_iterations_per_interval"] = w.max_newton, iters
        ta["jacobian_interval"] = 1
        cb05_oracle.run_batch(np.array([w.atol_newton, w.rtol_newton]), np.tile([w.atol_time, w.rtol_time], (m, 1)),
                              ta, ost, ot, odt, photo=w.photo)
        g = st.cpu().numpy()
        bad = np.nonzero(np.any(g != ost, axis=1) | (t.cpu().numpy() != ot) | (dt.cpu().numpy() != odt))[0]
        print("  cells differing:", bad.tolist())
        assert bad.size == 0, f"after {iters} iterations cells {bad.tolist()} differ from the oracle"


def test_persistent_schedule_is_bitwise_identical(fp64_vs_oracle, gpu, pkg):
    """The work-queue launch runs the same per-cell code: identical bits."""
    w, run, g_state, g_t, g_stats, *_ = fp64_vs_oracle
    run.reset()
    run.run_intervals(precision=pkg.PRECISION_FP64, schedule=pkg.SCHEDULE_PERSISTENT, lu=run.lu)
    gpu.cuda.synchronize()
    assert np.array_equal(run.state.cpu().numpy(), g_state)
    assert np.array_equal(run.t.cpu().numpy(), g_t)


@pytest.mark.parametrize("programs", ["lds", "memory"])
def test_program_placement_is_bitwise_identical(fp64_vs_oracle, programs, gpu, pkg):
    """The sparse form's compiled mechanism programs staged into LDS once per
    workgroup or read by every wavefront from device memory, with 1 and 4
    cells per workgroup: the same per-cell arithmetic, identical bits."""
    w, run, g_state, g_t, *_ = fp64_vs_oracle
    if run.lu != pkg.LU_SPARSE:
        pytest.skip("program placement applies to the sparse form")
    code = pkg.PROGRAMS_LDS if programs == "lds" else pkg.PROGRAMS_MEMORY
    for cpb in (1, 4):
        print("case: config0 seed 0, 64 cells, fp64 sparse, programs", programs, "cells per block", cpb, flush=True)
        run.reset()
        run.run_intervals(precision=pkg.PRECISION_FP64, schedule=pkg.SCHEDULE_GRID, lu=run.lu, cells_per_block=cpb,
                          programs=code)
        gpu.cuda.synchronize()
        assert np.array_equal(run.state.cpu().numpy(), g_state)
        assert np.array_equal(run.t.cpu().numpy(), g_t)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_finished_and_failed_cells_are_skipped(fp64_vs_oracle, precision, gpu, pkg):
    """runDeviceBatch skips a cell whose t_out already reaches t_end
    (src/core/TChem_AtmosphericChemistry.cpp:83) and, as the reference's
    failure convention implies, one whose dt_in is negative: state, t_out and
    dt_out stay as they were, the failure counter marks the failed ones.  The
    other cells of the batch give the same bits as in a launch without
    skipped cells; an empty batch is a no-op."""
    w, run, *_ = fp64_vs_oracle
    torch = gpu
    prec = pkg.PRECISION_FP64 if precision == "fp64" else pkg.PRECISION_MIXED
    n = w.n
    print("case: config0 seed 0, 64 cells, one interval,", precision, "lu", run.lu, flush=True)

    def one_interval(finished=(), failed=()):
        run.reset()
        ta = pkg.time_advance_from_tensor(run.tadv).copy()
        ta["_dt"][list(failed)] = -1.0
        run.tadv.copy_(torch.from_numpy(ta.view(np.uint8).copy()).to(run.tadv.device))
        run.t[list(finished)] = w.tend
        run.dt[list(finished) + list(failed)] = 5.0
        run.stats.zero_()
        pkg.AtmosphericChemistry.runDeviceBatch(run.tol_newton, run.tol_time, None, run.tadv, run.state, run.t,
                                                run.dt, run.state, run.kmcd, run.photo, run.stats, precision=prec,
                                                lu=run.lu)
        torch.cuda.synchronize()
        return (run.state.cpu().numpy(), run.t.cpu().numpy(), run.dt.cpu().numpy(),
                pkg.stats_from_tensor(run.stats))

    s0, t0, d0, _ = one_interval()
    finished, failed = list(range(0, 8)), list(range(8, 16))
    s1, t1, d1, st1 = one_interval(finished, failed)
    skipped = finished + failed
    live = list(range(16, n))
    np.testing.assert_array_equal(s1[skipped], w.state[skipped])
    assert np.all(t1[finished] == w.tend) and np.all(t1[failed] == w.tbeg)
    assert np.all(d1[skipped] == 5.0)
    assert np.all(st1["failed"][failed] == 1) and np.all(st1["failed"][finished] == 0)
    assert np.all(st1["accepted"][skipped] == 0) and np.all(st1["accepted"][live] > 0)
    np.testing.assert_array_equal(s1[live], s0[live])
    np.testing.assert_array_equal(t1[live], t0[live])
    np.testing.assert_array_equal(d1[live], d0[live])
    lib = pkg.lib()
    assert lib.tchem_atmchem_run_device_batch(pkg._Options(prec), 0, None, None, None, None, None, None, None, None,
                                              None, None, run.kmcd._h, None) == 0


@pytest.mark.parametrize("schedule", ["grid", "persistent"])
def test_captured_interval_loop_matches_eager_loop(fp64_vs_oracle, cb05, schedule, gpu, pkg):
    """The driver loop captured into a HIP graph (7 intervals per launch, the
    stop test on the device): after every replay the state equals the eager
    loop's state after the same number of intervals, bit for bit, with eager
    work and synchronisations between the replays; the loop stops after the
    same interval count."""
    torch = gpu
    w, run, g_state, g_t, *_ = fp64_vs_oracle
    kmd, kmcd, _ = cb05
    sched = pkg.SCHEDULE_GRID if schedule == "grid" else pkg.SCHEDULE_PERSISTENT
    k = 7
    print("case: captured loop,", schedule, "lu", run.lu, flush=True)
    run.reset()
    snaps, n_eager = [], 0
    while True:
        pkg.AtmosphericChemistry.runDeviceBatch(run.tol_newton, run.tol_time, None, run.tadv, run.state, run.t,
                                                run.dt, run.state, kmcd, run.photo, precision=pkg.PRECISION_FP64,
                                                schedule=sched, lu=run.lu)
        pkg.AtmosphericChemistry.updateTimeAdvance(run.tadv, run.t, run.dt)
        n_eager += 1
        stop = float(run.t.mean()) > w.tend * 0.9999
        if n_eager % k == 0 or stop:
            snaps.append(run.state.cpu().numpy().copy())
        if stop:
            break
    np.testing.assert_array_equal(snaps[-1], g_state)
    run.reset()
    graph = pkg.AtmosphericChemistry.captureGraph(run.tol_newton, run.tol_time, run.tadv, run.state, run.t, run.dt,
                                                  run.state, kmcd, k, 0.9999, photo_rates=run.photo,
                                                  precision=pkg.PRECISION_FP64, schedule=sched, lu=run.lu)
    other = torch.zeros(w.n, kmd.nActive, dtype=torch.float64, device="cuda")
    total, launches = 0, 0
    while True:
        graph.launch()
        n, stopped = graph.status()
        total += n
        launches += 1
        np.testing.assert_array_equal(run.state.cpu().numpy(), snaps[launches - 1])
        # eager work between replays
        pkg.SourceTerm.runDeviceBatch(run.state, run.photo, None, other, kmcd)
        torch.cuda.synchronize()
        if stopped:
            break
        assert n == k
    assert total == n_eager and launches == len(snaps)
    np.testing.assert_array_equal(run.t.cpu().numpy(), g_t)


def test_captured_graph_survives_kmcd_reuse(cb05, scen, drv, gpu, pkg):
    """A captured loop owns its rate constants and work-queue counter: a
    larger batch run on the same kmcd between capture and launch (which
    grows the kmcd's own rate-constant buffer) leaves the graph's results
    equal to the eager loop's, bit for bit; the kmcd refuses destruction
    while the graph exists and accepts it afterwards."""
    torch = gpu
    kmd, _, base = cb05
    kmcd = kmd.createConstData()  # a fresh kmcd: its rate-constant buffer starts small
    w = scen.config0(kmd, base[0], n=16, seed=4)
    print("case: captured graph, 16 cells, kmcd reused with 4096 cells before the launch", flush=True)
    run = drv.DeviceRun(kmcd, w)
    run.reset()
    run.run_intervals(precision=pkg.PRECISION_FP64, schedule=pkg.SCHEDULE_PERSISTENT, lu=pkg.LU_SPARSE)
    torch.cuda.synchronize()
    eager = run.state.cpu().numpy().copy()
    run.reset()
    graph = pkg.AtmosphericChemistry.captureGraph(run.tol_newton, run.tol_time, run.tadv, run.state, run.t, run.dt,
                                                  run.state, kmcd, 64, 0.9999, photo_rates=run.photo,
                                                  precision=pkg.PRECISION_FP64, schedule=pkg.SCHEDULE_PERSISTENT,
                                                  lu=pkg.LU_SPARSE)
    big = drv.DeviceRun(kmcd, scen.config0(kmd, base[0], n=4096, seed=5))
    big.reset()
    pkg.AtmosphericChemistry.runDeviceBatch(big.tol_newton, big.tol_time, None, big.tadv, big.state, big.t, big.dt,
                                            big.state, kmcd, big.photo, precision=pkg.PRECISION_FP64,
                                            schedule=pkg.SCHEDULE_PERSISTENT, lu=pkg.LU_SPARSE)
    torch.cuda.synchronize()
    for _ in range(1000):
        graph.launch()
        n, stopped = graph.status()
        if stopped or n < 64:
            break
    np.testing.assert_array_equal(run.state.cpu().numpy(), eager)
    lib = pkg.lib()
    assert lib.tchem_kmcd_destroy(kmcd._h) == -1  # TCHEM_ERR_INVALID
    assert "captured graph" in lib.tchem_last_error().decode()
    del graph
    assert lib.tchem_kmcd_destroy(kmcd._h) == 0
    kmcd._h = None


def test_autotuned_launch_shape_gives_identical_results(fp64_vs_oracle, cb05, gpu, pkg):
    """The autotuner returns a valid launch shape for the batch; the per-cell
    arithmetic does not depend on it (bitwise equal to the grid launch)."""
    w, run, g_state, g_t, *_ = fp64_vs_oracle
    kmcd = cb05[1]
    run.reset()
    tuned = pkg.AtmosphericChemistry.autotune(run.tol_newton, run.tol_time, run.tadv, run.state, kmcd,
                                              photo_rates=run.photo, precision=pkg.PRECISION_FP64,
                                              schedule=pkg.SCHEDULE_PERSISTENT, lu=run.lu)
    print("autotuned:", tuned, flush=True)
    assert tuned["lu"] == run.lu and tuned["cells_per_block"] >= 1 and tuned["blocks_per_cu"] >= 1
    assert tuned["seconds"] > 0
    np.testing.assert_array_equal(run.state.cpu().numpy(), w.state)  # the caller's arrays are untouched
    run.run_intervals(precision=pkg.PRECISION_FP64, schedule=pkg.SCHEDULE_PERSISTENT, lu=tuned["lu"],
                      blocks_per_cu=tuned["blocks_per_cu"], cells_per_block=tuned["cells_per_block"])
    gpu.cuda.synchronize()
    np.testing.assert_array_equal(run.state.cpu().numpy(), g_state)


def test_mixed_program_placement_is_bitwise_identical(fp64_vs_oracle, gpu, pkg):
    """Mixed sparse form: programs and coefficient table staged in LDS, or
    programs read from device memory with the coefficient table staged
    alone, at 4 and 16 cells per workgroup: the same per-cell arithmetic
    (FP32 pivot reciprocals included), identical bits."""
    w, run, *_ = fp64_vs_oracle
    if run.lu != pkg.LU_SPARSE:
        pytest.skip("program placement applies to the sparse form")
    out = {}
    for code, cpb in ((pkg.PROGRAMS_LDS, 4), (pkg.PROGRAMS_MEMORY, 4), (pkg.PROGRAMS_MEMORY, 16)):
        print("case: config0 seed 0, 64 cells, mixed sparse, programs", code, "cells per block", cpb, flush=True)
        run.reset()
        run.run_intervals(precision=pkg.PRECISION_MIXED, schedule=pkg.SCHEDULE_GRID, lu=run.lu, cells_per_block=cpb,
                          programs=code)
        gpu.cuda.synchronize()
        out[(code, cpb)] = (run.state.cpu().numpy(), run.t.cpu().numpy())
    ref = out[(pkg.PROGRAMS_LDS, 4)]
    assert np.all(ref[1] == w.tend)
    for k, (s, t) in out.items():
        assert np.array_equal(s, ref[0]), k
        assert np.array_equal(t, ref[1]), k


def test_mixed_precision_error_against_fp64_oracle(fp64_vs_oracle, cb05, gpu, pkg):
    """FP32 iteration matrix and LU, FP64 state/residual/error/step control,
    against the FP64 oracle (64 config-0 cells, 3600 s): the relative error
    per species is at most 1e-4 wherever the integrator controls relative
    accuracy — species above atol / 1e-4 = 1e-8 ppm, below which the
    absolute tolerance dominates — and the tolerance-normalised error is
    below 1 everywhere.  The floor-1e-12 maximum is printed for the record:
    two FP64 solvers that differ only in the elimination order of their LU
    already differ by up to 4.7e-4 there on 1024 cells, through accept /
    reject and Newton-count decisions at their thresholds."""
    w, run, g_state, g_t, g_stats, o_state, o_t, o_stats = fp64_vs_oracle
    m = cb05[0].nActive
    run.reset()
    run.run_to_end(precision=pkg.PRECISION_MIXED, schedule=pkg.SCHEDULE_PERSISTENT, collect_stats=True, lu=run.lu)
    gpu.cuda.synchronize()
    assert np.all(run.t.cpu().numpy() == w.tend)
    assert run.total_stats[:, 5].sum() == 0
    a, b = run.state.cpu().numpy()[:, 3:3 + m], o_state[:, 3:3 + m]
    e8 = relerr(a, b, 1e-8)
    e12 = relerr(a, b, 1e-12)
    tn = np.abs(a - b) / (w.atol_time + w.rtol_time * np.abs(b))
    print("mixed vs oracle: max", e8.max(), "(floor 1e-8)", e12.max(), "(floor 1e-12), median", np.median(e12),
          "tolerance-normalised max", tn.max())
    assert e8.max() <= 1e-4
    assert tn.max() < 1.0
    assert np.median(e12) < 1e-10


@pytest.mark.parametrize("lu", LU_FORMS)
def test_forced_newton_failure_on_gpu(cb05, gpu, pkg, lu):
    """One Newton iteration and a minimum step of 100 s: every cell fails the
    reference way (species zeroed, t_out = tend, dt_out = -1) and the kernel exits."""
    kmd, kmcd, base = cb05
    print("case: forced Newton failure, lu", lu, flush=True)
    torch = gpu
    m = kmd.nActive
    n = 4
    st = torch.tensor(np.repeat(base, n, 0), device="cuda")
    tn = torch.tensor([1e-30, 1e-14], dtype=torch.float64, device="cuda")
    tt = torch.tensor(np.tile([1e-12, 1e-6], (m, 1)), device="cuda")
    tadv = pkg.time_advance_tensor(n, 0.0, 3600.0, 100.0, 100.0, 100.0, 1, 5, 1)
    t = torch.zeros(n, dtype=torch.float64, device="cuda")
    dt = torch.zeros(n, dtype=torch.float64, device="cuda")
    stats = pkg.stats_tensor(n)
    pkg.AtmosphericChemistry.runDeviceBatch(tn, tt, None, tadv, st, t, dt, st, kmcd, stats=stats,
                                            lu=lu_code(pkg, lu))
    torch.cuda.synchronize()
    s = pkg.stats_from_tensor(stats)
    assert np.all(s["failed"] == 1) and np.all(s["newton_failures"] == 4)
    assert np.all(t.cpu().numpy() == 3600.0) and np.all(dt.cpu().numpy() == -1.0)
    out = st.cpu().numpy()
    assert np.all(out[:, 3:3 + m] == 0.0)
    # a second launch skips the failed cells
    pkg.AtmosphericChemistry.runDeviceBatch(tn, tt, None, tadv, st, t, dt, st, kmcd, lu=lu_code(pkg, lu))
    torch.cuda.synchronize()
    assert np.array_equal(st.cpu().numpy(), out)


def test_invalid_arguments_are_rejected(cb05, gpu, pkg):
    kmd, kmcd, base = cb05
    torch = gpu
    st = torch.tensor(base, device="cuda")
    with pytest.raises(pkg.TChemError, match="device tensor"):
        pkg.SourceTerm.runDeviceBatch(torch.tensor(base), None, None, torch.zeros(1, 74), kmcd)
    with pytest.raises(pkg.TChemError, match="float64"):
        pkg.SourceTerm.runDeviceBatch(st.float(), None, None, torch.zeros(1, 74, device="cuda"), kmcd)
    lib = pkg.lib()
    opts = pkg._Options(5, 0, 0)
    assert lib.tchem_atmchem_run_device_batch(opts, 1, None, None, None, None, None, None, None, None, None, None,
                                              kmcd._h, None) == -1
    assert "precision" in lib.tchem_last_error().decode()
    opts = pkg._Options(0, 0, 0, 7)
    assert lib.tchem_atmchem_run_device_batch(opts, 1, None, None, None, None, None, None, None, None, None, None,
                                              kmcd._h, None) == -1
    assert "lu" in lib.tchem_last_error().decode()
    opts = pkg._Options()
    opts.reserved[1] = 1
    assert lib.tchem_atmchem_run_device_batch(opts, 1, None, None, None, None, None, None, None, None, None, None,
                                              kmcd._h, None) == -1
    assert "reserved" in lib.tchem_last_error().decode()
    opts = pkg._Options(lu=pkg.LU_DENSE, cells_per_block=2)
    assert lib.tchem_atmchem_run_device_batch(opts, 1, None, None, None, None, None, None, None, None, None, None,
                                              kmcd._h, None) == -1
    assert "one cell per block" in lib.tchem_last_error().decode()
    opts = pkg._Options(lu=pkg.LU_SPARSE, cells_per_block=64)
    assert lib.tchem_atmchem_run_device_batch(opts, 1, None, None, None, None, None, None, None, None, None, None,
                                              kmcd._h, None) == -1
    assert "cells_per_block" in lib.tchem_last_error().decode()
    # workspace of each form: one cell of the dense form; the sparse form's
    # shared program plus its cells
    dense = pkg.AtmosphericChemistry.getWorkSpaceSize(kmcd, pkg.PRECISION_MIXED, pkg.LU_DENSE)
    one = pkg.AtmosphericChemistry.getWorkSpaceSize(kmcd, pkg.PRECISION_MIXED, pkg.LU_SPARSE, 1)
    two = pkg.AtmosphericChemistry.getWorkSpaceSize(kmcd, pkg.PRECISION_MIXED, pkg.LU_SPARSE, 2)
    full = pkg.AtmosphericChemistry.getWorkSpaceSize(kmcd, pkg.PRECISION_MIXED, pkg.LU_SPARSE)
    cell = two - one
    cells = round((full - (one - cell)) / cell)  # one - cell: the programs' share (0 when read from memory)
    assert cell < dense / 2 and full <= 160 * 1024 and (full + cell > 160 * 1024 or cells == 16)
    opts = pkg._Options(lu=pkg.LU_SPARSE, programs=5)
    assert lib.tchem_atmchem_run_device_batch(opts, 1, None, None, None, None, None, None, None, None, None, None,
                                              kmcd._h, None) == -1
    assert "programs" in lib.tchem_last_error().decode()
