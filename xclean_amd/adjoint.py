"""The differentiable window solve: the Gauss-Newton step of every window as a torch function of the observed pixel
coordinates and per-observation weights, with the implicit-function adjoint as its backward.

Forward: the step dx = -(H + 1e-9 I)^-1 b at the solver's current state (vio/window.py:933-941 at lambda = 0; the
state is normally the LM solution, so dx is the last, converged update), factored in FP64 on the GPU.  Backward:
libxba ``xba_backward`` reuses that Cholesky factor for y = F^-1 g and forms dL/dz and dL/dw per observation row
in one kernel (the formulas are at the top of csrc/xba_adj.hip).  There is no CPU path.
"""
from __future__ import annotations

import torch

from .window import BatchSolver


class WindowStep(torch.autograd.Function):
    """dx (sum n,) = GN step of every window of ``solver`` at its current state, as a function of z (M, 2) and
    w (M,), both in the concatenated original observation order of ``solver.problems``."""

    @staticmethod
    def forward(ctx, z, w, solver):
        solver.set_measurements(z.detach(), w.detach())
        dx = solver.gn_step().clone()
        ctx.solver = solver
        return dx

    @staticmethod
    def backward(ctx, g):
        gz, gw = ctx.solver.adjoint(g.detach().to(torch.float64))
        return gz, gw, None


def solve_differentiable(problems, z, w, iters=10, lam0=1e-4, precision="mixed", **kw):
    """LM-solve the windows (``iters`` iterations from their state, mixed precision by default), then return the
    converged Gauss-Newton step as a differentiable function of z and w, and the solver (whose state is the
    solution).  z (M, 2) and w (M,) are tensors in the concatenated original observation order."""
    solver = BatchSolver(problems, precision=precision, with_weights=True, **kw)
    solver.set_measurements(z.detach(), w.detach())
    solver.upload(lam0)
    solver.iterate(iters)
    return WindowStep.apply(z, w, solver), solver
