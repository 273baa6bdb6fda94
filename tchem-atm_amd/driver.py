"""Box-model driver: the host interval loop of TChem_AtmosphericChemistry.x.

Restates src/examples/TChem_AtmosphericChemistry.cpp:262-352 over the C ABI:
relaunch runDeviceBatch until the mean cell time reaches 0.9999 * tend, and
between launches carry t_out/dt_out into tadv._tbeg/_dt.  The persistent
form integrates every cell to tend in one launch instead.
"""
from __future__ import annotations

import importlib

import numpy as np

T = importlib.import_module("tchem-atm_amd")


class DeviceRun:
    """Device buffers of one workload, reusable across repeated runs."""

    def __init__(self, kmcd, workload, device="cuda"):
        import torch
        self.torch = torch
        self.kmcd = kmcd
        self.w = workload
        m = kmcd.nActive
        n = workload.n
        self.n = n
        tn = np.array([workload.atol_newton, workload.rtol_newton])
        tt = np.tile([workload.atol_time, workload.rtol_time], (m, 1))
        self.tol_newton = torch.tensor(tn, dtype=torch.float64, device=device)
        self.tol_time = torch.tensor(tt, dtype=torch.float64, device=device)
        self.state0 = torch.tensor(workload.state, dtype=torch.float64, device=device)
        self.photo = torch.tensor(workload.photo, dtype=torch.float64, device=device) if kmcd.nPhoto > 0 else None
        self.state = torch.empty_like(self.state0)
        self.t = torch.empty(n, dtype=torch.float64, device=device)
        self.dt = torch.empty(n, dtype=torch.float64, device=device)
        self.stats = T.stats_tensor(n, device)
        self.device = device

    def reset(self, iters_per_interval=None):
        w = self.w
        self.state.copy_(self.state0)
        self.t.fill_(w.tbeg)
        self.dt.fill_(w.dtmax)  # examples/...:227 (not read: tadv._dt is what the integrator uses)
        ipi = w.iters_per_interval if iters_per_interval is None else iters_per_interval
        tadv = T.time_advance_tensor(self.n, w.tbeg, w.tend, w.dt_in, w.dtmin, w.dtmax, w.max_newton, ipi,
                                     w.jacobian_interval, self.device)
        if getattr(self, "tadv", None) is not None and self.tadv.shape == tadv.shape:
            self.tadv.copy_(tadv)  # in place: a captured graph holds this buffer
        else:
            self.tadv = tadv

    def run_intervals(self, precision=T.PRECISION_FP64, schedule=T.SCHEDULE_GRID, max_intervals=100000,
                      blocks_per_cu=0, collect_stats=False, stream=None, lu=T.LU_AUTO, cells_per_block=0,
                      programs=T.PROGRAMS_AUTO):
        """Reference host loop; returns the number of launches."""
        torch = self.torch
        tend = self.w.tend
        it = 0
        total = None
        while it < max_intervals:
            T.AtmosphericChemistry.runDeviceBatch(self.tol_newton, self.tol_time, None, self.tadv, self.state,
                                                  self.t, self.dt, self.state, self.kmcd, self.photo,
                                                  self.stats if collect_stats else None, precision=precision,
                                                  schedule=schedule, blocks_per_cu=blocks_per_cu, lu=lu,
                                                  cells_per_block=cells_per_block, programs=programs,
                                                  stream=stream)
            if collect_stats:
                s = T.stats_from_tensor(self.stats)
                arr = np.stack([s[f] for f in T.COUNTERS], 1).astype(np.int64)
                arr[:, 5] = s["failed"]
                if total is None:
                    total = arr.copy()
                else:
                    total[:, :5] += arr[:, :5]
                    total[:, 5] = np.maximum(total[:, 5], arr[:, 5])
            T.AtmosphericChemistry.updateTimeAdvance(self.tadv, self.t, self.dt, stream)
            it += 1
            if float(self.t.mean()) > tend * 0.9999:
                break
        self.total_stats = total
        return it

    def run_intervals_graph(self, precision=T.PRECISION_FP64, schedule=T.SCHEDULE_GRID, intervals_per_launch=32,
                            max_intervals=100000, blocks_per_cu=0, lu=T.LU_AUTO, stream=None):
        """The same loop as run_intervals, captured: each graph launch runs
        up to intervals_per_launch intervals with the stop test on the
        device; the host reads the stop flag once per launch.  Returns the
        number of intervals run (equal to run_intervals')."""
        key = (precision, schedule, intervals_per_launch, blocks_per_cu, lu)
        if getattr(self, "_graph_key", None) != key:
            self.graph = T.AtmosphericChemistry.captureGraph(
                self.tol_newton, self.tol_time, self.tadv, self.state, self.t, self.dt, self.state, self.kmcd,
                intervals_per_launch, 0.9999, photo_rates=self.photo, precision=precision, schedule=schedule,
                blocks_per_cu=blocks_per_cu, lu=lu)
            self._graph_key = key
        total = 0
        while total < max_intervals:
            self.graph.launch(stream)
            n, stopped = self.graph.status()
            total += n
            if stopped or n < intervals_per_launch:
                break
        return total

    def start_slices(self, n_slices=1, ends=None):
        """Resets the run for run_slice: every cell at tbeg with dt_in, the
        interval [tbeg, tend] cut into n_slices equal slices, or into slices
        ending at the given times (the last one tend)."""
        w = self.w
        self.reset()
        if ends is None:
            ends = [w.tbeg + (w.tend - w.tbeg) * (k + 1) / n_slices for k in range(n_slices)]
            ends[-1] = w.tend
        self.slice_ends = [float(e) for e in ends]
        self.tadv = T.time_advance_tensor(self.n, w.tbeg, self.slice_ends[0], w.dt_in, w.dtmin, w.dtmax,
                                          w.max_newton, 10_000_000, w.jacobian_interval, self.device)
        self._tadv64 = self.tadv.view(self.torch.float64).view(self.n, 7)
        self._graph_key = None
        self.acc_stats = self.torch.zeros((self.n, 6), dtype=self.torch.int64, device=self.device)

    def run_slice(self, k, precision=T.PRECISION_MIXED, schedule=T.SCHEDULE_PERSISTENT, blocks_per_cu=0,
                  lu=T.LU_AUTO, cells_per_block=0, programs=T.PROGRAMS_AUTO, collect_stats=True, stream=None):
        """Slice k of the hour: every cell from its own t_out to the slice end
        in one launch (persistent: one work-queue launch), then t_out / dt_out
        carried into the next slice's TimeAdvance records.  Cells that
        failed keep dt = -1 and stay failed.  Per-cell counters accumulate
        in acc_stats (accepted, rejected, Newton failures, Newton
        iterations, factorisations, failed)."""
        torch = self.torch
        if k > 0:
            self._tadv64[:, 0].copy_(self.t)
            self._tadv64[:, 1].fill_(self.slice_ends[k])
            self._tadv64[:, 2].copy_(self.dt)
        T.AtmosphericChemistry.runDeviceBatch(self.tol_newton, self.tol_time, None, self.tadv, self.state, self.t,
                                              self.dt, self.state, self.kmcd, self.photo,
                                              self.stats if collect_stats else None, precision=precision,
                                              schedule=schedule, blocks_per_cu=blocks_per_cu, lu=lu,
                                              cells_per_block=cells_per_block, programs=programs, stream=stream)
        if collect_stats:
            s = self.stats.view(torch.int32).view(self.n, 10)
            self.acc_stats[:, :5] += s[:, :5]
            self.acc_stats[:, 5] = torch.maximum(self.acc_stats[:, 5], s[:, 5].to(torch.int64))

    def run_to_end(self, precision=T.PRECISION_MIXED, schedule=T.SCHEDULE_PERSISTENT, blocks_per_cu=0,
                   collect_stats=False, stream=None, lu=T.LU_AUTO, cells_per_block=0, programs=T.PROGRAMS_AUTO):
        """Every cell to tend in one launch (iteration cap 10^7 per cell)."""
        w = self.w
        self.tadv = T.time_advance_tensor(self.n, w.tbeg, w.tend, w.dt_in, w.dtmin, w.dtmax, w.max_newton,
                                          10_000_000, w.jacobian_interval, self.device)
        self._graph_key = None  # a captured loop holds the previous tadv buffer
        T.AtmosphericChemistry.runDeviceBatch(self.tol_newton, self.tol_time, None, self.tadv, self.state, self.t,
                                              self.dt, self.state, self.kmcd, self.photo,
                                              self.stats if collect_stats else None, precision=precision,
                                              schedule=schedule, blocks_per_cu=blocks_per_cu, lu=lu,
                                              cells_per_block=cells_per_block, programs=programs, stream=stream)
        if collect_stats:
            s = T.stats_from_tensor(self.stats)
            self.total_stats = np.stack([s[f] for f in T.COUNTERS], 1).astype(np.int64)
            self.t_start, self.t_end = s["t_start"].copy(), s["t_end"].copy()
        return 1

    def run_slice_intervals(self, k, iters=10, precision=T.PRECISION_FP64, schedule=T.SCHEDULE_GRID, lu=T.LU_DENSE,
                            collect_stats=True, stream=None, max_intervals=1_000_000):
        """Slice k through the reference's host interval loop: launches of at
        most `iters` TR-BDF2 iterations, t_out / dt_out carried into tadv
        after each, until every cell reached the slice end (a failed cell
        counts as done: its t_out is the slice end).  The iteration cap does
        not change a cell's step sequence, so the result equals run_slice's.
        Returns the number of launches."""
        torch = self.torch
        end = self.slice_ends[k]
        if k > 0:
            self._tadv64[:, 0].copy_(self.t)
            self._tadv64[:, 1].fill_(end)
            self._tadv64[:, 2].copy_(self.dt)
        ta32 = self.tadv.view(torch.int32).view(self.n, 14)
        ta32[:, 11].fill_(iters)  # _num_time_iterations_per_interval
        launches = 0
        while launches < max_intervals:
            T.AtmosphericChemistry.runDeviceBatch(self.tol_newton, self.tol_time, None, self.tadv, self.state,
                                                  self.t, self.dt, self.state, self.kmcd, self.photo,
                                                  self.stats if collect_stats else None, precision=precision,
                                                  schedule=schedule, lu=lu, stream=stream)
            if collect_stats:
                s = self.stats.view(torch.int32).view(self.n, 10)
                self.acc_stats[:, :5] += s[:, :5]
                self.acc_stats[:, 5] = torch.maximum(self.acc_stats[:, 5], s[:, 5].to(torch.int64))
            T.AtmosphericChemistry.updateTimeAdvance(self.tadv, self.t, self.dt, stream)
            launches += 1
            if bool((self.t >= end).all()):
                break
        ta32[:, 11].fill_(10_000_000)
        return launches
