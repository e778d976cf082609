"""torch.autograd bindings of the RK4 step and of the c^2 misfit.

    RK4Step.apply(q, solver, dt)                 -> S q   (backward: S^T g)
    MisfitC2.apply(c2, solver, q0, nsteps, dt, data[, checkpoint_every]) -> J
                                                  (backward: dJ/dc2 by the
                                                  checkpointed discrete adjoint)

Tensors are float64 of the library's shapes, q: (6, 5, nr, n, n), c2:
(6, nr, n, n); they may live on the CPU or the GPU (they are staged through
host memory: these wrappers compose the library's transposes with torch code,
they are not the timed path).  The compute is the HIP library; there is no
fallback.
"""
import numpy as np
import torch

__all__ = ["RK4Step", "MisfitC2"]


def _np(t):
    return t.detach().to("cpu", torch.float64).contiguous().numpy()


class RK4Step(torch.autograd.Function):
    """One RK4 step (timestep.f90:41-86) as a differentiable map of the state."""

    @staticmethod
    def forward(ctx, q, solver, dt):
        ctx.solver, ctx.dt = solver, float(dt)
        solver.set_state(_np(q))
        solver.rk4(1, dt)
        return torch.from_numpy(solver.get_state()).to(q.device)

    @staticmethod
    def backward(ctx, g):
        out = ctx.solver.adj_step(_np(g), ctx.dt)
        return torch.from_numpy(out).to(g.device), None, None


class MisfitC2(torch.autograd.Function):
    """J(c2) = sum of squared p' residuals at the solver's receivers over
    nsteps RK4 steps from q0 with the c^2 field c2."""

    @staticmethod
    def forward(ctx, c2, solver, q0, nsteps, dt, data, checkpoint_every=0):
        solver.set_c2_field(_np(c2))
        J, grad, _ = solver.misfit_gradient(_np(q0), int(nsteps), float(dt), np.asarray(data, dtype=np.float64),
                                            checkpoint_every)
        ctx.save_for_backward(torch.from_numpy(grad).to(c2.device))
        return torch.tensor(J, dtype=torch.float64, device=c2.device)

    @staticmethod
    def backward(ctx, gJ):
        (grad,) = ctx.saved_tensors
        return gJ * grad, None, None, None, None, None, None
