"""The MFMA form: the dense FP32 LU on the matrix cores.

- tchem_dense_lu_solve_mfma against numpy's FP64 solve on seeded diagonally
  dominant systems of every size class the 32 x 32 tiling distinguishes (one
  tile, a tile boundary, a partly filled last tile, the full 255), and the
  singular-pivot convention (NaN for that system only).  A wrong register layout,
  tile owner or panel orientation gives O(1) errors; FP32 factors of
  well-conditioned systems give about 1e-6.
- The TR-BDF2 integrator with lu = TCHEM_LU_MFMA on the uniformly coupled
  synthetic 250-species mechanism (its LU factor is 48 % dense in the
  fill-reducing order) against the FP64 oracle, to the mixed form's bars:
  tolerance-normalised error below 1 and 1e-4 relative above 1e-8 ppm.
"""
import ctypes
import dataclasses
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _systems(n, batch, seed):
    g = np.random.Generator(np.random.Philox(key=seed))
    A = g.uniform(-1.0, 1.0, (batch, n, n)) / np.sqrt(n)
    A[:, np.arange(n), np.arange(n)] += 2.0 + g.uniform(0.0, 1.0, (batch, n))
    b = g.uniform(-1.0, 1.0, (batch, n))
    return A, b


def _solve(pkg, gpu, A, b):
    torch = gpu
    lib = pkg.lib()
    n, batch = A.shape[1], A.shape[0]
    dA = torch.tensor(A, dtype=torch.float64, device="cuda")
    db = torch.tensor(b, dtype=torch.float64, device="cuda")
    dx = torch.full((batch, n), -7.0, dtype=torch.float64, device="cuda")
    rc = lib.tchem_dense_lu_solve_mfma(n, batch, ctypes.c_void_p(dA.data_ptr()), ctypes.c_void_p(db.data_ptr()),
                                       ctypes.c_void_p(dx.data_ptr()), ctypes.c_void_p(pkg._stream_ptr(None)))
    assert rc == 0, lib.tchem_last_error().decode()
    torch.cuda.synchronize()
    return dx.cpu().numpy()


@pytest.mark.parametrize("n", [1, 7, 32, 33, 100, 224, 250, 255])
def test_dense_lu_solve_mfma_matches_fp64_solve(pkg, gpu, n):
    print("case: dense MFMA solve n =", n, flush=True)
    A, b = _systems(n, 6, 100 + n)
    x = _solve(pkg, gpu, A, b)
    ref = np.linalg.solve(A, b[..., None])[..., 0]
    err = np.abs(x - ref).max(axis=1) / np.abs(ref).max(axis=1)
    print("  max-norm relative error per system", err.tolist())
    assert err.max() < 2e-5


def test_dense_lu_solve_mfma_zero_pivot(pkg, gpu):
    A, b = _systems(40, 3, 7)
    A[1, 0, :2] = 0.0  # no interchanges: the first 2 x 2 pivot block of system 1 is singular
    x = _solve(pkg, gpu, A, b)
    assert np.isnan(x[1]).all()
    ref = np.linalg.solve(A[[0, 2]], b[[0, 2], :, None])[..., 0]
    assert np.abs(x[[0, 2]] - ref).max() < 2e-5


def test_dense_lu_solve_mfma_sizes_rejected(pkg, gpu):
    lib = pkg.lib()
    assert lib.tchem_dense_lu_solve_mfma(256, 1, None, None, None, None) != 0
    assert lib.tchem_dense_lu_solve_mfma(0, 1, None, None, None, None) != 0
    assert lib.tchem_dense_lu_solve_mfma(10, 0, None, None, None, None) == 0


@pytest.fixture(scope="module")
def uniform(gpu, pkg, oracle_mod, tmp_path_factory):
    S = importlib.import_module("tchem-atm_amd.scenarios")
    mech = S.synthetic_mechanism_yaml(str(tmp_path_factory.mktemp("mech") / "uniform250.yaml"), coupling="uniform")
    kmd = pkg.KineticModelData(mech)
    kmcd = kmd.createConstData()
    base = pkg.setScenarioConditions(mech, kmd)
    w, _ = S.config4(kmd, base[0], size_scale=1 / 64)
    w = dataclasses.replace(w, state=w.state[:8].copy(), photo=w.photo[:8].copy(), tend=3600.0)
    o = oracle_mod.Oracle(mech)
    n, m = w.n, o.m
    st = w.state.copy()
    t = np.zeros(n)
    dt = np.full(n, w.dtmax)
    ta = np.zeros(n, oracle_mod.TADV_DTYPE)
    ta["tbeg"], ta["tend"], ta["dt"], ta["dtmin"], ta["dtmax"] = 0.0, w.tend, w.dt_in, w.dtmin, w.dtmax
    ta["max_num_newton_iterations"] = w.max_newton
    ta["num_time_iterations_per_interval"] = 10_000_000
    ta["jacobian_interval"] = 1
    s = np.zeros(n, oracle_mod.STATS_DTYPE)
    o.run_batch(np.array([w.atol_newton, w.rtol_newton]), np.tile([w.atol_time, w.rtol_time], (m, 1)), ta, st, t,
                dt, photo=w.photo, stats=s, nthreads=8)
    return kmd, kmcd, w, st, s


@pytest.mark.parametrize("schedule", ["grid", "persistent"])
def test_mfma_integrator_against_oracle(uniform, gpu, pkg, schedule):
    kmd, kmcd, w, o_state, o_stats = uniform
    D = importlib.import_module("tchem-atm_amd.driver")
    m = kmd.nActive
    print("case: MFMA LU form, uniform synthetic 250-species mechanism, 8 cells, 3600 s,", schedule, flush=True)
    run = D.DeviceRun(kmcd, w)
    run.reset()
    sch = pkg.SCHEDULE_GRID if schedule == "grid" else pkg.SCHEDULE_PERSISTENT
    run.run_to_end(precision=pkg.PRECISION_MIXED, schedule=sch, collect_stats=True, lu=pkg.LU_MFMA)
    gpu.cuda.synchronize()
    st = run.total_stats
    assert np.all(run.t.cpu().numpy() == w.tend)
    assert st[:, 5].sum() == 0 and o_stats["failed"].sum() == 0
    g = run.state.cpu().numpy()[:, 3:3 + m]
    b = o_state[:, 3:3 + m]
    e8 = np.abs(g - b) / np.maximum(np.abs(b), 1e-8)
    tn = np.abs(g - b) / (w.atol_time + w.rtol_time * np.abs(b))
    print("  steps gpu", st[:, 0].tolist(), "oracle", o_stats["accepted"].tolist())
    print("  Newton its gpu", st[:, 3].tolist(), "oracle", o_stats["newton_iterations"].tolist())
    print("  rel err above 1e-8 ppm", e8.max(axis=1).tolist(), "tolerance-normalised max", tn.max())
    assert tn.max() < 1.0
    assert e8.max() < 1e-4


def test_mfma_integrator_matches_mixed_sparse_steps(uniform, gpu, pkg):
    """The MFMA form and the mixed sparse form factor the same FP32 matrix in
    the same pivot order, so their step counts agree closely and their end
    states agree to the mixed form's accuracy."""
    kmd, kmcd, w, _, _ = uniform
    D = importlib.import_module("tchem-atm_amd.driver")
    m = kmd.nActive
    out = {}
    for lu in (pkg.LU_MFMA, pkg.LU_SPARSE):
        run = D.DeviceRun(kmcd, w)
        run.reset()
        run.run_to_end(precision=pkg.PRECISION_MIXED, schedule=pkg.SCHEDULE_PERSISTENT, collect_stats=True, lu=lu)
        gpu.cuda.synchronize()
        out[lu] = (run.state.cpu().numpy()[:, 3:3 + m], run.total_stats.copy())
    (a, sa), (b, sb) = out[pkg.LU_MFMA], out[pkg.LU_SPARSE]
    print("  steps mfma", sa[:, 0].tolist(), "sparse", sb[:, 0].tolist())
    assert np.all(np.abs(sa[:, 0] - sb[:, 0]) <= 0.02 * sb[:, 0] + 2)
    e8 = np.abs(a - b) / np.maximum(np.abs(b), 1e-8)
    assert e8.max() < 1e-4
