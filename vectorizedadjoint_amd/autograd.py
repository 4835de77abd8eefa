"""PyTorch autograd over the device forward / adjoint pair.

    yT = odeint_adjoint(solver, y0, p)      # y0 [N, B], p [P, B] float64 device tensors
    loss(yT).backward()                     # y0.grad = dJ/dy0, p.grad = dJ/dp

The forward is the library's forward sweep (va_forward); the backward is its
discrete-adjoint reverse sweep (va_adjoint) seeded with dJ/dy(T), i.e. the exact
gradient of the computed, step-size-frozen map y0, p -> y(T), as the reference's
adjointSolve returns it (lib/include/detail/backpropagation.hpp:280-348).
Instances whose forward failed (status != 0) get NaN gradients.

The reverse sweep differentiates the trajectory stored by the LAST forward on
the solver, so a second forward on the same solver before backward() raises.
Use one Solver per differentiated call that is still pending.
"""
from __future__ import annotations

import torch

from . import Solver


class OdeAdjointFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, y0, p, solver: Solver):
        y = y0.detach().clone(memory_format=torch.contiguous_format)
        pc = p.detach().contiguous()
        solver.forward(y, pc)
        ctx.solver = solver
        ctx.generation = solver.generation
        ctx.save_for_backward(pc)
        return y

    @staticmethod
    def backward(ctx, grad_yT):
        solver = ctx.solver
        if solver.generation != ctx.generation:
            raise RuntimeError("odeint_adjoint: the solver ran another forward before backward(); "
                               "its stored trajectory belongs to that later call")
        (p,) = ctx.saved_tensors
        N, P, B = solver.problem.n_state, solver.problem.n_param, solver.problem.batch
        lam = grad_yT.detach().to(torch.float64).contiguous().clone().reshape(1, N, B)
        mu = torch.zeros((1, P, B), dtype=torch.float64, device=p.device)
        solver.adjoint(p, lam, mu)
        return lam[0], mu[0], None


def odeint_adjoint(solver: Solver, y0: torch.Tensor, p: torch.Tensor) -> torch.Tensor:
    """y(T) of every instance, differentiable with respect to y0 and p."""
    return OdeAdjointFunction.apply(y0, p, solver)
