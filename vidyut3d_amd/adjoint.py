"""torch.autograd binding of the electrode-current objective and its
hand-written discrete adjoint (vd_objective_forward / vd_adjoint_gradient).

    theta = torch.tensor(sim.get_params(), dtype=torch.float64, requires_grad=True)
    J = electrode_current(theta, sim, nsteps=400, checkpoint_every=20)
    J.backward()            # theta.grad = dJ/dtheta from the adjoint kernels

theta holds the case's run-time parameters followed by the voltage amplitude
of the electrode of current location 0.  The forward re-initialises the state
(initdomaindata) and runs nsteps steps; the backward recomputes each
checkpointed segment on the GPU and runs the adjoint steps backwards.  torch
only carries theta and the gradient; no torch operation touches the state.
A backward must follow its own forward: a second forward on the same
simulation replaces the checkpoints.
"""
from __future__ import annotations

import numpy as np
import torch


class ElectrodeCurrent(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, sim, nsteps, checkpoint_every):
        p = theta.detach().to("cpu", torch.float64).numpy()
        sim.set_params(p)
        sim.init_state()
        J = sim.objective_forward(int(nsteps), int(checkpoint_every))
        ctx.sim = sim
        return torch.tensor(J, dtype=theta.dtype, device=theta.device)

    @staticmethod
    def backward(ctx, grad_J):
        g = ctx.sim.adjoint_gradient()
        gt = torch.from_numpy(np.ascontiguousarray(g)).to(grad_J.device, grad_J.dtype)
        return grad_J * gt, None, None, None


def electrode_current(theta: torch.Tensor, sim, nsteps: int, checkpoint_every: int = 20) -> torch.Tensor:
    """Time-integrated current through the objective electrode after nsteps
    steps from the initial state, differentiable in theta."""
    return ElectrodeCurrent.apply(theta, sim, nsteps, checkpoint_every)
