"""The no-allocation entry point, tchem_atmchem_run_device_batch_noalloc,
against tchem_atmchem_run_device_batch through the C ABI: its rate
constants and work-queue counter live in the library's file-scope device
scratch array instead of library allocations.  Two calls in a row, the
second with a smaller batch, for the sparse persistent and the dense grid
forms: the same bits as the allocating entry point.  A batch beyond the
scratch limit is rejected before any device work."""
import ctypes
import importlib

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup(gpu, pkg):
    kmd = pkg.KineticModelData(golden("cb05cl_ae5.yaml"))
    kmcd = kmd.createConstData()
    base = pkg.setScenarioConditions(golden("cb05cl_ae5.yaml"), kmd)
    S = importlib.import_module("tchem-atm_amd.scenarios")
    lib = pkg.lib()
    fn = lib.tchem_atmchem_run_device_batch_noalloc
    fn.argtypes = lib.tchem_atmchem_run_device_batch.argtypes
    fn.restype = ctypes.c_int
    return kmd, kmcd, base, S, lib, fn


def one_call(pkg, torch, fn, kmcd, w, opts, iters):
    """One runDeviceBatch interval of `iters` TR-BDF2 iterations through
    `fn`; returns state, t_out, dt_out, counters."""
    m = kmcd.nActive
    n = w.n
    dev = "cuda"
    tn = torch.tensor([w.atol_newton, w.rtol_newton], dtype=torch.float64, device=dev)
    tt = torch.tensor(np.tile([w.atol_time, w.rtol_time], (m, 1)), device=dev)
    tadv = pkg.time_advance_tensor(n, w.tbeg, w.tend, w.dt_in, w.dtmin, w.dtmax, w.max_newton, iters, 1, dev)
    st = torch.tensor(w.state, device=dev)
    ph = torch.tensor(w.photo, device=dev)
    t = torch.zeros(n, dtype=torch.float64, device=dev)
    dt = torch.zeros(n, dtype=torch.float64, device=dev)
    stats = pkg.stats_tensor(n, dev)
    rc = fn(ctypes.byref(opts), n, tn.data_ptr(), tt.data_ptr(), None, tadv.data_ptr(), st.data_ptr(),
            ph.data_ptr(), t.data_ptr(), dt.data_ptr(), st.data_ptr(), stats.data_ptr(), kmcd._h,
            torch.cuda.current_stream().cuda_stream)
    assert rc == 0, pkg.lib().tchem_last_error().decode()
    torch.cuda.synchronize()
    s = pkg.stats_from_tensor(stats)
    return st.cpu().numpy(), t.cpu().numpy(), dt.cpu().numpy(), np.stack([s[f] for f in pkg.COUNTERS], 1)


@pytest.mark.parametrize("form", ["sparse-persistent", "dense-grid"])
def test_noalloc_matches_allocating_entry_point(setup, gpu, pkg, form):
    kmd, kmcd, base, S, lib, fn = setup
    torch = gpu
    if form == "sparse-persistent":
        opts = pkg._Options(pkg.PRECISION_MIXED, pkg.SCHEDULE_PERSISTENT, 0, pkg.LU_SPARSE)
    else:
        opts = pkg._Options(pkg.PRECISION_FP64, pkg.SCHEDULE_GRID, 0, pkg.LU_DENSE)
    for n, iters in ((512, 400), (96, 300)):  # the second call: a smaller batch
        print("case: no-alloc", form, n, "cells,", iters, "iterations", flush=True)
        w = S.config1(kmd, base[0], seed=1, cells=np.arange(0, 131072, 131072 // n)[:n])
        got = one_call(pkg, torch, fn, kmcd, w, opts, iters)
        ref = one_call(pkg, torch, lib.tchem_atmchem_run_device_batch, kmcd, w, opts, iters)
        for g, r in zip(got, ref):
            np.testing.assert_array_equal(g, r)
        assert np.all(got[1] > 0.0)


def test_noalloc_rejects_batches_beyond_the_scratch(setup, gpu, pkg):
    """One cell more than the 16 MiB scratch holds: an error naming the
    limit, returned on the host (the arrays are valid, nothing runs)."""
    kmd, kmcd, base, S, lib, fn = setup
    torch = gpu
    opts = pkg._Options(pkg.PRECISION_MIXED, pkg.SCHEDULE_PERSISTENT, 0, pkg.LU_SPARSE)
    cap = ((16 << 20) - 256) // (8 * kmd.nReac)
    n = cap + 1
    w = S.config1(kmd, base[0], seed=1, cells=np.arange(n))
    st = torch.tensor(w.state, device="cuda")
    before = st.clone()
    tn = torch.tensor([1e-18, 1e-8], dtype=torch.float64, device="cuda")
    tt = torch.tensor(np.tile([1e-12, 1e-6], (kmd.nActive, 1)), device="cuda")
    tadv = pkg.time_advance_tensor(n, 0.0, 10.0, 1e-20, 1e-20, 60.0, 20, 1, 1)
    t = torch.zeros(n, dtype=torch.float64, device="cuda")
    dt = torch.zeros(n, dtype=torch.float64, device="cuda")
    rc = fn(ctypes.byref(opts), n, tn.data_ptr(), tt.data_ptr(), None, tadv.data_ptr(), st.data_ptr(), None,
            t.data_ptr(), dt.data_ptr(), st.data_ptr(), None, kmcd._h, None)
    assert rc == -1
    msg = lib.tchem_last_error().decode()
    assert "scratch" in msg and f"at most {cap} cells" in msg, msg
    torch.cuda.synchronize()
    assert torch.equal(st, before) and torch.all(t == 0.0)
