"""torch.autograd binding of the design gradient: J(gamma) as a differentiable function.

Forward runs lbmx_gradient (forward sweeps with checkpoints, reverse sweeps with
recomputation) and keeps dJ/dgamma; backward scales it by the incoming gradient. PyTorch is
only the tensor front end here: the computation is the HIP engine behind the C ABI. A design
tensor on the GPU stays there: it is handed to the engine by device pointer and dJ/dgamma is
copied device to device from the engine's own array.
"""
import numpy as np
import torch

import lbmx


class DesignObjective(torch.autograd.Function):
    """J = objective of `solver` after `nsteps` steps from rest, as a function of the design
    field gamma (natural node order, float64). Usage: DesignObjective.apply(gamma, solver, nsteps)."""

    @staticmethod
    def forward(ctx, gamma, solver, nsteps, rho0=1.0):
        if gamma.is_cuda:
            solver.set_design_device(gamma.detach().to(torch.float64).contiguous())
            J, _ = solver.gradient(int(nsteps), rho0=rho0, want_sens=False)
            sens = solver.sensitivity_device().clone().to(gamma.dtype)
        else:
            g = gamma.detach().to(torch.float64).contiguous().numpy()
            solver.set_design(g)
            J, sens = solver.gradient(int(nsteps), rho0=rho0)
            sens = torch.from_numpy(sens).to(gamma.dtype)
        ctx.save_for_backward(sens)
        return gamma.new_tensor(J, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_out):
        (sens,) = ctx.saved_tensors
        return grad_out * sens, None, None, None


def objective(gamma, solver, nsteps, rho0=1.0):
    return DesignObjective.apply(gamma, solver, nsteps, rho0)


def central_difference(solver, gamma, nsteps, cells, h=1e-5):
    """Central finite differences of J with respect to gamma at the given node indices."""
    g = np.asarray(gamma, dtype=np.float64).ravel().copy()
    out = []
    for c in cells:
        gp, gm = g.copy(), g.copy()
        gp[c] += h
        gm[c] -= h
        solver.set_design(gp)
        solver.init_equilibrium()
        Jp = solver.forward(nsteps)
        solver.set_design(gm)
        solver.init_equilibrium()
        Jm = solver.forward(nsteps)
        out.append((Jp - Jm) / (2 * h))
    solver.set_design(g)
    return np.array(out)
