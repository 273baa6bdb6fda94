"""Single-particle SIMPOL.1 aerosol path on the GPU (C ABI:
tchem_aerosol_rhs_run_device_batch, tchem_aerosolchem_run_device_batch)
against the CPU oracle and the reference's goldens.

Error metric per unknown: |gpu - oracle| / max(|oracle|, floor), floor = the
integration's absolute tolerance."""
import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu

CASES = {
    "simpol": ("simpol_gas.yaml", "simpol_aero.yaml", "simpol_particles.yaml"),
    "soa": ("cb05_soa_gas.yaml", "cb05_soa_aero.yaml", "cb05_soa_particles.yaml"),
}


@pytest.fixture(scope="module")
def models(gpu, pkg, oracle_mod):
    out = {}
    for name, (chem, aero, part) in CASES.items():
        kmd = pkg.KineticModelData(golden(chem))
        kmcd = kmd.createConstData()
        amd = pkg.AerosolModelData(golden(aero), kmd)
        amcd = amd.createConstData(kmcd)
        o = oracle_mod.AerosolOracle(golden(chem), golden(aero))
        out[name] = (kmd, kmcd, amd, amcd, o)
    return out


def scenario(pkg, kmd, amd, name, n):
    chem, _, part = CASES[name]
    g = pkg.setScenarioConditions(golden(chem), kmd)
    state = np.zeros((1, amd.stateDim))
    state[:, :g.shape[1]] = g[:1]
    nc = amd.scenarioConditionParticles(golden(part), state)
    return np.repeat(state, n, axis=0), np.repeat(nc, n, axis=0)


def random_cells(pkg, kmd, amd, name, n, seed):
    """Scenario state with perturbed T, p, gas species, particle loads and
    number concentrations."""
    rng = np.random.default_rng(seed)
    state, nc = scenario(pkg, kmd, amd, name, n)
    state[:, 1] = rng.uniform(7e4, 1.01e5, n)
    state[:, 2] = rng.uniform(260.0, 300.0, n)
    ma = kmd.nActive
    state[:, 3:3 + ma] *= 10.0 ** rng.uniform(-0.5, 0.5, (n, ma))
    state[:, 3:3 + ma] += 10.0 ** rng.uniform(-6, -3, (n, ma))  # SOA precursors start non-zero
    nc = 10.0 ** rng.uniform(4, 7, (n, amd.nParticles))
    base = 3 + kmd.nSpec
    k = amd.nSpec * amd.nParticles
    state[:, base:base + k] = 10.0 ** rng.uniform(-16, -12, (n, k))
    return state, nc


def test_product_loader_matches_oracle_reader(models):
    for name, (kmd, kmcd, amd, amcd, o) in models.items():
        assert amd.speciesNames == o.a.names and amd.nParticles == o.a.nParticles
        np.testing.assert_array_equal(amd.density, o.a.density)
        np.testing.assert_array_equal(amd.molecularWeights, o.a.mw)
        assert amd.nUnknowns == o.mt and amd.stateDim == o.dim


def test_simpol_rhs_golden_on_gpu(models, gpu, pkg):
    """TChem_Test_Simpol_RHS (Simpol_RHS_HOST.txt, 6 printed digits)."""
    torch = gpu
    kmd, kmcd, amd, amcd, o = models["simpol"]
    state = np.zeros((1, amd.stateDim))
    state[0, 1], state[0, 2] = 101253.3, 272.5
    state[0, 3] = 0.1
    state[0, 3 + kmd.nSpec] = 1.0e-8 / 1.3e6
    state[0, 4 + kmd.nSpec] = 1.4e-2 / 1.3e6
    om = torch.full((1, amd.nUnknowns), np.nan, dtype=torch.float64, device="cuda")
    nc = torch.tensor([[1.3e6]], dtype=torch.float64, device="cuda")
    pkg.AerosolRHS.runDeviceBatch(torch.tensor(state, device="cuda"), nc, om, amcd)
    got = om.cpu().numpy()[0]
    ref = np.array([float(ln.split()[1]) for ln in open(golden("simpol_rhs_host.txt")) if not ln.startswith("#")])
    np.testing.assert_array_equal(np.array([float(f"{v:e}") for v in got]), ref)


@pytest.mark.parametrize("name", list(CASES))
def test_rhs_matches_oracle(models, gpu, pkg, name):
    torch = gpu
    kmd, kmcd, amd, amcd, o = models[name]
    state, nc = random_cells(pkg, kmd, amd, name, 256, seed=11)
    om = torch.full((256, amd.nUnknowns), np.nan, dtype=torch.float64, device="cuda")
    pkg.AerosolRHS.runDeviceBatch(torch.tensor(state, device="cuda"), torch.tensor(nc, device="cuda"), om, amcd)
    got = om.cpu().numpy()
    ref = o.rhs_batch(state, nc)
    scale = np.abs(ref).max(axis=1, keepdims=True)
    assert (np.abs(got - ref) / scale).max() < 1e-13
    nz = np.abs(ref) > 1e-30 * scale
    assert (np.abs(got - ref)[nz] / np.abs(ref)[nz]).max() < 1e-9


def gpu_intervals(pkg, torch, amcd, amd, state, nc, tend, dtmax, tol_newton, atol_time, rtol_time, steps,
                  max_intervals, dtmin=1e-20, max_newton=100, rows=False, lu=None):
    """The TChem_AerosolChemistry.x interval loop on the GPU: per interval one
    runDeviceBatch, then tadv <- (t_out, dt_out)."""
    n = state.shape[0]
    st = torch.tensor(state, device="cuda")
    ncd = torch.tensor(nc, device="cuda")
    tadv = pkg.time_advance_tensor(n, 0.0, tend, dtmin, dtmin, dtmax, max_newton, steps)
    t = torch.zeros(n, dtype=torch.float64, device="cuda")
    dt = torch.zeros(n, dtype=torch.float64, device="cuda")
    tn = torch.tensor(tol_newton, dtype=torch.float64, device="cuda")
    tt = torch.tensor(np.tile([atol_time, rtol_time], (amd.nUnknowns, 1)), device="cuda")
    stats = pkg.stats_tensor(n)
    total = np.zeros((n, 6), np.int64)
    out = [np.c_[np.full(n, -1.0), np.zeros(n), np.zeros(n), state]]
    for it in range(max_intervals):
        pkg.AerosolChemistry.runDeviceBatch(tn, tt, None, tadv, st, ncd, t, dt, st, amcd, stats=stats,
                                            lu=pkg.LU_AUTO if lu is None else lu)
        pkg.AtmosphericChemistry.updateTimeAdvance(tadv, t, dt)
        s = pkg.stats_from_tensor(stats)
        for j, f in enumerate(pkg.COUNTERS):
            total[:, j] += s[f]
        tt_host = t.cpu().numpy()
        if rows:
            out.append(np.c_[np.full(n, float(it)), tt_host, dt.cpu().numpy(), st.cpu().numpy()])
        if tt_host.mean() > tend * 0.9999:
            break
    return st.cpu().numpy(), t.cpu().numpy(), dt.cpu().numpy(), total, (np.vstack(out) if rows else None)


def oracle_intervals(oracle_mod, o, state, nc, tend, dtmax, tol_newton, atol_time, rtol_time, steps, max_intervals,
                     dtmin=1e-20, max_newton=100):
    n = state.shape[0]
    st = state.copy()
    tadv = np.zeros(n, oracle_mod.TADV_DTYPE)
    tadv["tend"], tadv["dt"], tadv["dtmin"], tadv["dtmax"] = tend, dtmin, dtmin, dtmax
    tadv["max_num_newton_iterations"], tadv["num_time_iterations_per_interval"] = max_newton, steps
    tadv["jacobian_interval"] = 1
    t, dt = np.zeros(n), np.zeros(n)
    total = np.zeros(n, oracle_mod.STATS_DTYPE)
    tt = np.tile([atol_time, rtol_time], (o.mt, 1))
    for _ in range(max_intervals):
        s = np.zeros(n, oracle_mod.STATS_DTYPE)
        o.run_batch(np.array(tol_newton), tt, tadv, st, nc, t, dt, stats=s)
        for f in ("accepted", "rejected", "newton_failures", "newton_iterations", "factorizations"):
            total[f] += s[f]
        total["failed"] |= s["failed"]
        tadv["tbeg"], tadv["dt"] = t, dt
        if t.mean() > tend * 0.9999:
            break
    return st, t, dt, total


@pytest.mark.parametrize("lu", ["sparse", "dense"])
@pytest.mark.parametrize("ref_name,name,cells", [
    ("simpol_full_gas.npy", "simpol", 1),
    ("simpol_batched_full_gas.npy", "simpol", 3),
    ("cb05_soa_batched_full_gas.npy", "soa", 3),
])
def test_tines_ode_goldens_on_gpu(models, gpu, pkg, ref_name, name, cells, lu):
    """TinesODE ctest (dtmin 1e-20, dtmax 10, tend 600, Newton (1e-18, 1e-8),
    two intervals of 10 steps, time tolerances (1e-12, 1e-4)) with
    compare_tchem.py's norm criterion at 1e-8."""
    kmd, kmcd, amd, amcd, o = models[name]
    state, nc = scenario(pkg, kmd, amd, name, cells)
    print("case:", ref_name, lu, flush=True)
    *_, rows = gpu_intervals(pkg, gpu, amcd, amd, state, nc, 600.0, 10.0, (1e-18, 1e-8), 1e-12, 1e-4, 10, 2,
                             rows=True, lu=pkg.LU_SPARSE if lu == "sparse" else pkg.LU_DENSE)
    # rows are interval-major like the .dat writer
    ref = np.load(golden(ref_name))
    assert rows.shape == ref.shape
    diff = rows - ref
    ref_l2 = np.linalg.norm(ref, 2, 0)
    denom = np.where(ref_l2 > 1e-23, ref_l2, 1.0)
    worst = max((np.linalg.norm(diff, k, 0) / denom).max() for k in (1, 2, np.inf))
    assert worst <= 1e-8, worst


@pytest.fixture(scope="module")
def soa_oracle_run(models, pkg, oracle_mod):
    kmd, kmcd, amd, amcd, o = models["soa"]
    state, nc = random_cells(pkg, kmd, amd, "soa", 48, seed=5)
    return state, nc, oracle_intervals(oracle_mod, o, state, nc, *SOA_ARGS)


SOA_ARGS = (600.0, 10.0, (1e-12, 1e-8), 1e-12, 1e-6, 100, 1000)


@pytest.mark.parametrize("lu", ["sparse", "dense"])
def test_integration_matches_oracle(models, gpu, pkg, soa_oracle_run, lu):
    """CB05CL_AE5 + SOA partitioning, 48 perturbed cells, 600 s: the FP64 GPU
    forms (sparse programs, dense LU) against the oracle.  Cells with
    identical step sequences agree to 1e-9; any other cell stays within the
    integration tolerance."""
    kmd, kmcd, amd, amcd, o = models["soa"]
    n = 48
    state, nc, (o_state, o_t, o_dt, o_stats) = soa_oracle_run
    print("case: soa 48 cells seed 5", lu, flush=True)
    g_state, g_t, g_dt, g_stats, _ = gpu_intervals(pkg, gpu, amcd, amd, state, nc, *SOA_ARGS,
                                                   lu=pkg.LU_SPARSE if lu == "sparse" else pkg.LU_DENSE)
    assert np.all(g_t == 600.0) and np.all(o_t == 600.0)
    assert g_stats[:, 5].sum() == 0 and o_stats["failed"].sum() == 0
    cols = np.r_[3:3 + kmd.nActive, 3 + kmd.nSpec:amd.stateDim]
    errs = np.abs(g_state[:, cols] - o_state[:, cols]) / np.maximum(np.abs(o_state[:, cols]), 1e-12)
    # particle species are kg m-3 per particle (~1e-18): their floor is relative to the particle mass
    part = np.abs(g_state[:, 3 + kmd.nSpec:] - o_state[:, 3 + kmd.nSpec:]) / np.maximum(
        np.abs(o_state[:, 3 + kmd.nSpec:]), 1e-30)
    err = np.maximum(errs[:, :kmd.nActive].max(axis=1), part.max(axis=1))
    same = (g_stats[:, 0] == o_stats["accepted"]) & (g_stats[:, 1] == o_stats["rejected"])
    print("identical step sequences:", same.sum(), "of", n, "max err there:", err[same].max(),
          "elsewhere:", err[~same].max() if (~same).any() else None)
    assert same.sum() >= 0.8 * n
    assert err[same].max() < 1e-9
    assert err.max() < 1e-3
    # partitioning did something: every cell's SOA mass moved
    base = 3 + kmd.nSpec
    assert np.all(np.abs(g_state[:, base + 1:base + 5] - state[:, base + 1:base + 5]).max(axis=1) > 0)
    np.testing.assert_array_equal(g_state[:, :3], state[:, :3])
    np.testing.assert_array_equal(g_state[:, 3 + kmd.nActive:3 + kmd.nSpec], state[:, 3 + kmd.nActive:3 + kmd.nSpec])


def test_driver_golden_on_gpu(models, gpu, pkg):
    """TChem_Driver ctest state, one 30 s step under its solver file, on the
    GPU integrator; compare_files.py's np.isclose against output_ref.txt."""
    kmd, kmcd, amd, amcd, o = models["soa"]
    s = np.zeros((1, amd.stateDim))
    s[0, 1], s[0, 2] = 81060.0, 270.5
    s[0, 50], s[0, 51], s[0, 36], s[0, 81] = 0.06, 1e-06, 0.001, 1e8
    s[0, 3 + kmd.nSpec] = 1e-08
    nc = np.full((1, amd.nParticles), 10000.0)
    g_state, g_t, *_ = gpu_intervals(pkg, gpu, amcd, amd, s, nc, 30.0, 30.0, (1e-6, 1e-4), 1e-12, 1e-6, 200000, 1,
                                     dtmin=1e-12)
    assert g_t[0] == 30.0
    ref = np.loadtxt(golden("driver_output_ref.txt"))
    out = np.array([float(f"{v:e}") for v in g_state[0]])
    assert np.all(np.isclose(out, ref)), np.argwhere(~np.isclose(out, ref)).ravel()


def test_invalid_arguments_are_rejected(models, gpu, pkg):
    torch = gpu
    kmd, kmcd, amd, amcd, o = models["simpol"]
    st = torch.zeros((2, amd.stateDim), dtype=torch.float64, device="cuda")
    with pytest.raises(pkg.TChemError, match="null device array"):
        pkg.AerosolRHS.runDeviceBatch(st, None, st, amcd)
    opts = pkg._Options(pkg.PRECISION_MIXED, 0, 0, 0, 0, 0)
    import ctypes
    rc = pkg.lib().tchem_aerosolchem_run_device_batch(ctypes.byref(opts), 2, None, None, None, None, None, None,
                                                      None, None, None, None, amcd._h, None)
    assert rc == -3 and b"FP64 only" in pkg.lib().tchem_last_error()


def test_tchem_driver_ctest_through_the_executable(tmp_path, gpu, pkg):
    """The reference's DriverInterface ctest on TChem_Driver.x (the coupling
    interface, libtchem_driver.so): gas species names and aerosol properties
    line by line, the state after one 30 s step with np.isclose
    (src/unit-tests/TChem_Driver/compare_files.py)."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(pkg.LIB_PATH), "TChem_Driver.x")
    print("case: TChem_Driver.x cb05_soa", flush=True)
    r = subprocess.run([exe, golden("cb05_soa_gas.yaml"), golden("cb05_soa_aero.yaml"),
                        golden("driver_config_solver.yaml")], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    for mine, ref in (("gas_species_names.txt", "driver_gas_species_names_ref.txt"),
                      ("aero_species_props.txt", "driver_aero_species_props_ref.txt")):
        assert [ln.strip() for ln in open(tmp_path / mine)] == [ln.strip() for ln in open(golden(ref))]
    out = np.loadtxt(tmp_path / "output.txt")
    ref = np.loadtxt(golden("driver_output_ref.txt"))
    assert out.shape == ref.shape and np.all(np.isclose(out, ref))


def test_driver_interface_integrates_every_batch_entry(gpu, pkg):
    """Two batch entries through libtchem_driver.so: entry 0 is the ctest
    state (matches output_ref.txt), entry 1 the same at 280 K; both advance
    (the reference copies back entry 0 only)."""
    import ctypes
    import os
    lib = ctypes.CDLL(os.path.join(os.path.dirname(pkg.LIB_PATH), "libtchem_driver.so"))
    lib.TChem_getLastError.restype = ctypes.c_char_p
    lib.initialize(golden("cb05_soa_gas.yaml").encode(), golden("cb05_soa_aero.yaml").encode(),
                   golden("driver_config_solver.yaml").encode(), 2)
    assert lib.TChem_getLastError() == b""
    n = lib.TChem_getLengthOfStateVector()
    n_gas = lib.TChem_getNumberOfSpecies()
    rows = np.zeros((2, n))
    rows[:, 1], rows[:, 2] = 81060.0, 270.5
    rows[1, 2] = 280.0
    rows[:, 50], rows[:, 51], rows[:, 36], rows[:, 81] = 0.06, 1e-06, 0.001, 1e8
    rows[:, 3 + n_gas] = 1e-08
    nc = np.full(lib.TChem_getNumberConcentrationVectorSize(), 10000.0)
    for i in range(2):
        r = np.ascontiguousarray(rows[i])
        lib.TChem_setStateVector(r.ctypes.data_as(ctypes.c_void_p), i)
        lib.TChem_setNumberConcentrationVector(nc.ctypes.data_as(ctypes.c_void_p), i)
    dt = ctypes.c_double(30.0)
    lib.TChem_doTimestep(ctypes.byref(dt))
    assert lib.TChem_getLastError() == b""
    out = np.zeros((2, n))
    lib.TChem_getAllStateVectorHost(out.ctypes.data_as(ctypes.c_void_p))
    lib.finalize()
    ref = np.loadtxt(golden("driver_output_ref.txt"))
    assert np.all(np.isclose(np.array([float(f"{v:e}") for v in out[0]]), ref))
    moved = np.abs(out[1] - rows[1]) > 1e-6 * np.abs(rows[1])
    assert moved[3:3 + n_gas].sum() > 10 and np.all(out[1, :3] == rows[1, :3])
    assert np.abs(out[1] - out[0])[3:].max() > 0


def test_many_particles_match_oracle(gpu, pkg, oracle_mod):
    """20 particles per cell (178 unknowns: the sparse form's 4-slot
    instance; the dense form is limited to 128 unknowns): 8 perturbed cells
    with different particle loads per particle, 300 s, against the oracle."""
    from oracle import amd as amd_mod
    kmd = pkg.KineticModelData(golden("cb05_soa_gas.yaml"))
    kmcd = kmd.createConstData()
    amd = pkg.AerosolModelData(golden("cb05_soa_aero.yaml"), kmd)
    amd.setNumberofParticles(20)
    amcd = amd.createConstData(kmcd)
    assert amd.nUnknowns == kmd.nActive + 5 * 20
    o = oracle_mod.AerosolOracle(golden("cb05_soa_gas.yaml"), golden("cb05_soa_aero.yaml"))
    a = o.a
    a.nParticles = 20
    o = oracle_mod.AerosolOracle(o.g, a)
    rng = np.random.default_rng(21)
    n = 8
    g = pkg.setScenarioConditions(golden("cb05_soa_gas.yaml"), kmd)
    state = np.zeros((n, amd.stateDim))
    state[:, :g.shape[1]] = g[0]
    state[:, 2] = rng.uniform(265, 295, n)
    ma = kmd.nActive
    state[:, 3:3 + ma] += 10.0 ** rng.uniform(-6, -3, (n, ma))
    base = 3 + kmd.nSpec
    state[:, base:] = 10.0 ** rng.uniform(-17, -13, (n, 100))
    nc = 10.0 ** rng.uniform(4, 6, (n, 20))
    args = (300.0, 10.0, (1e-12, 1e-8), 1e-12, 1e-6, 100, 1000)
    print("case: soa 20 particles 8 cells", flush=True)
    g_state, g_t, _, g_stats, _ = gpu_intervals(pkg, gpu, amcd, amd, state, nc, *args)
    o_state, o_t, _, o_stats = oracle_intervals(oracle_mod, o, state, nc, *args)
    assert np.all(g_t == 300.0) and g_stats[:, 5].sum() == 0 and o_stats["failed"].sum() == 0
    same = (g_stats[:, 0] == o_stats["accepted"]) & (g_stats[:, 1] == o_stats["rejected"])
    gas = np.abs(g_state[:, 3:3 + ma] - o_state[:, 3:3 + ma]) / np.maximum(np.abs(o_state[:, 3:3 + ma]), 1e-12)
    part = np.abs(g_state[:, base:] - o_state[:, base:]) / np.maximum(np.abs(o_state[:, base:]), 1e-30)
    err = np.maximum(gas.max(axis=1), part.max(axis=1))
    print("identical step sequences:", same.sum(), "of", n, "max err there:", err[same].max())
    assert same.sum() >= 6 and err[same].max() < 1e-9 and err.max() < 1e-3
    # the dense form refuses more than 128 unknowns
    st = gpu.tensor(state, device="cuda")
    with pytest.raises(pkg.TChemError, match="at most 128"):
        pkg.AerosolChemistry.runDeviceBatch(gpu.zeros(2, dtype=gpu.float64, device="cuda"),
                                            gpu.zeros((amd.nUnknowns, 2), dtype=gpu.float64, device="cuda"),
                                            None, pkg.time_advance_tensor(n, 0, 1, 1e-6, 1e-6, 1, 10, 1), st,
                                            gpu.tensor(nc, device="cuda"), gpu.zeros(n, dtype=gpu.float64,
                                                                                     device="cuda"),
                                            gpu.zeros(n, dtype=gpu.float64, device="cuda"), st, amcd,
                                            lu=pkg.LU_DENSE)
