"""torch.autograd.Function over the discrete-adjoint C ABI.

    problem = AdjointProblem(kmcd, tol_newton, tol_time, tbeg=0.0, tend=1800.0, ...)
    final = KineticsAdjoint.apply(log_rate_scale, state, problem)   # [n][nSpec+3]
    loss = (w * torch.log(final[:, 3:3 + m] + 1e-12)).sum()
    loss.backward()   # log_rate_scale.grad = dL/dtheta (summed over cells
                      # for a shared [nReac] theta, per cell for [n][nReac]),
                      # state.grad[:, 3:] = dL/d(initial species)

The forward pass is the FP64 sparse persistent integrator with its accepted
steps recorded (checkpoint every 16); the backward pass is one launch of the
adjoint kernel (segment replay + transposed stage solves).  Rate constants
are kf_r = k_r exp(theta_r), so for photolysis reactions theta is a uniform
log-scale on the per-cell rate.  Pressure and temperature are inputs without
gradient; gradients of failed cells are NaN.
"""
from __future__ import annotations

import importlib

import torch

T = importlib.import_module("tchem-atm_amd")


class AdjointProblem:
    """Everything but the differentiable inputs: mechanism, tolerances, the
    time window and the per-cell photolysis rates."""

    def __init__(self, kmcd, tol_newton, tol_time, tbeg, tend, dt=1e-20, dtmin=1e-20, dtmax=60.0, max_newton=20,
                 jacobian_interval=1, max_steps=4096, max_iterations=1 << 30, photo_rates=None):
        self.kmcd = kmcd
        self.tol_newton, self.tol_time = tol_newton, tol_time
        self.tbeg, self.tend, self.dt, self.dtmin, self.dtmax = tbeg, tend, dt, dtmin, dtmax
        self.max_newton, self.jacobian_interval = max_newton, jacobian_interval
        self.max_steps, self.max_iterations = max_steps, max_iterations
        self.photo_rates = photo_rates


class KineticsAdjoint(torch.autograd.Function):
    @staticmethod
    def forward(ctx, log_rate_scale, state, problem):
        p = problem
        n = state.shape[0]
        dev = state.device
        out = state.detach().clone()
        t = torch.full((n,), float(p.tbeg), dtype=torch.float64, device=dev)
        dt = torch.zeros(n, dtype=torch.float64, device=dev)
        tadv = T.time_advance_tensor(n, p.tbeg, p.tend, p.dt, p.dtmin, p.dtmax, p.max_newton,
                                     min(p.max_iterations, 2 ** 31 - 1), p.jacobian_interval, dev)
        ws = T.Adjoint.workspace(p.kmcd, n, p.max_steps, dev)
        theta = log_rate_scale.detach().contiguous()
        T.Adjoint.forward(p.tol_newton, p.tol_time, tadv, out, t, dt, ws, p.max_steps, p.kmcd,
                          photo_rates=p.photo_rates, log_rate_scale=theta)
        ctx.problem = p
        ctx.per_cell = theta.dim() == 2
        ctx.ws = ws
        ctx.n = n
        return out

    @staticmethod
    def backward(ctx, grad_out):
        p = ctx.problem
        lam = grad_out[:, 3:].contiguous().to(torch.float64)
        lam_out = torch.empty_like(lam)
        grad = torch.empty((ctx.n, p.kmcd.nReac), dtype=torch.float64, device=lam.device)
        T.Adjoint.backward(lam, lam_out, grad, ctx.ws, p.max_steps, p.kmcd)
        grad_state = torch.zeros_like(grad_out)
        grad_state[:, 3:] = lam_out
        g = grad if ctx.per_cell else grad.sum(0)
        return g, grad_state, None
