"""PyTorch harness for gradient checks of the residual: a torch.autograd.Function whose
forward is mose_residual and whose backward is the hand-written HIP discrete adjoint
(mose_residual_vjp).  Tensors are CPU float64 arrays in the reference's blk%P / blk%R
layout, one per block; the solver's device state is overwritten by the forward."""
import numpy as np
import torch


class ResidualFunction(torch.autograd.Function):
    """R_1..R_nb = residual(P_1..P_nb) on a GpuSolver; backward = discrete adjoint."""

    @staticmethod
    def forward(ctx, solver, *P):
        ctx.solver = solver
        for b, Pb in enumerate(P, start=1):
            solver.put_state(b, Pb.detach().cpu().numpy())
        ctx.P = [Pb.detach().cpu().numpy().copy() for Pb in P]
        out = tuple(torch.from_numpy(solver.residual(b)) for b in range(1, len(P) + 1))
        return out if len(out) > 1 else out[0]

    @staticmethod
    def backward(ctx, *Rbar):
        solver = ctx.solver
        for b, Pb in enumerate(ctx.P, start=1):      # the adjoint is taken at the forward state
            solver.put_state(b, Pb)
        Rb = [np.zeros_like(ctx.P[b]) if g is None else g.detach().cpu().numpy() for b, g in enumerate(Rbar)]
        Pbar = solver.residual_vjp(Rb)
        return (None,) + tuple(torch.from_numpy(x) for x in Pbar)


def residual(solver, *P):
    return ResidualFunction.apply(solver, *P)
