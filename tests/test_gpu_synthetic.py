"""The cascade variant of the synthetic 250-species mechanism (seeded, NCAR
YAML through the same loader; its LU factor stays sparse): its LU /
evaluation programs do not fit in LDS, so the sparse form reads them from
device memory.  GPU against the FP64 oracle on 8 cells of the
uneven-batch workload over 60 s (about 1700 steps per cell)."""
import dataclasses
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def synth(gpu, pkg, oracle_mod, tmp_path_factory):
    S = importlib.import_module("tchem-atm_amd.scenarios")
    mech = S.synthetic_mechanism_yaml(str(tmp_path_factory.mktemp("mech") / "synthetic250.yaml"), coupling="cascade")
    kmd = pkg.KineticModelData(mech)
    kmcd = kmd.createConstData()
    base = pkg.setScenarioConditions(mech, kmd)
    w, _ = S.config4(kmd, base[0], size_scale=1 / 64)
    w = dataclasses.replace(w, state=w.state[:8].copy(), photo=w.photo[:8].copy(), tend=60.0)
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
                dt, photo=w.photo, stats=s)
    return kmd, kmcd, w, st, s


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_synthetic_mechanism_against_oracle(synth, gpu, pkg, precision):
    kmd, kmcd, w, o_state, o_stats = synth
    D = importlib.import_module("tchem-atm_amd.driver")
    m = kmd.nActive
    print("case: synthetic 250-species mechanism, 8 cells, 60 s,", precision, flush=True)
    run = D.DeviceRun(kmcd, w)
    run.reset()
    prec = pkg.PRECISION_FP64 if precision == "fp64" else pkg.PRECISION_MIXED
    run.run_to_end(precision=prec, schedule=pkg.SCHEDULE_PERSISTENT, collect_stats=True, lu=pkg.LU_SPARSE)
    gpu.cuda.synchronize()
    st = run.total_stats
    assert np.all(run.t.cpu().numpy() == w.tend)
    assert st[:, 5].sum() == 0 and o_stats["failed"].sum() == 0
    g = run.state.cpu().numpy()[:, 3:3 + m]
    err = np.abs(g - o_state[:, 3:3 + m]) / np.maximum(np.abs(o_state[:, 3:3 + m]), 1e-12)
    same = (st[:, 0] == o_stats["accepted"]) & (st[:, 1] == o_stats["rejected"])
    print("steps gpu", st[:, 0].tolist(), "oracle", o_stats["accepted"].tolist(), "max err", err.max(axis=1).tolist())
    if precision == "fp64":
        assert same.sum() >= 6
        assert err[same].max() < 1e-8
    assert np.median(err) < 1e-8
    assert err.max() < 1e-2
