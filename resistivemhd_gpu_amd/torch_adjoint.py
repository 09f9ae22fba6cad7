"""torch.autograd.Function over the discrete adjoint of the RK2 step.

    UN = rollout(U0, eta, cfgs, nsteps)

U0 is a float64 tensor (B, nx, ny, 8) of initial states in the reference's
AoS order (ghosts included; the boundary is applied before the first step),
eta a float64 tensor (B, nx, ny) of per-cell resistivities, cfgs B
configurations of the same grid.  The forward runs the HIP library's batched
step with every dt recorded and sqrt(nsteps) checkpoints; backward runs the
hand-written discrete adjoint (mhd_adjoint_backward) with those dt frozen and
returns dL/dU0 and dL/deta.  Compose U0 from primitive fields with torch ops
to get gradients with respect to the initial velocity.
"""
import numpy as np
import torch

from .mhd2d import ResistiveMHD2DAdjoint


class MHDRollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, U0, eta, cfgs, nsteps, stride):
        if U0.dtype != torch.float64 or eta.dtype != torch.float64:
            raise TypeError("the step is FP64: pass float64 tensors")
        B = U0.shape[0]
        if len(cfgs) != B or eta.shape[0] != B:
            raise ValueError("one configuration and one eta field per run")
        adj = ResistiveMHD2DAdjoint(cfgs)
        U0n = U0.detach().to("cpu").numpy()
        etan = eta.detach().to("cpu").numpy()
        for r in range(B):
            adj.setU0(r, U0n[r])
            adj.setEtaField(r, etan[r])
        adj.forward(int(nsteps), int(stride))
        UN = np.stack([adj.getU(r) for r in range(B)])
        ctx.adj = adj
        ctx.device = U0.device
        return torch.from_numpy(UN).to(U0.device)

    @staticmethod
    def backward(ctx, gUN):
        adj = ctx.adj
        g = gUN.detach().to("cpu").contiguous().numpy()
        for r in range(len(adj)):
            adj.set_seed(r, g[r])
        adj.backward()
        gU0 = np.stack([adj.grad_state(r) for r in range(len(adj))])
        geta = np.stack([adj.grad_eta(r) for r in range(len(adj))])
        return torch.from_numpy(gU0).to(ctx.device), torch.from_numpy(geta).to(ctx.device), None, None, None


def rollout(U0, eta, cfgs, nsteps, stride=0):
    """Final states after nsteps RK2 steps, differentiable in U0 and eta."""
    return MHDRollout.apply(U0, eta, cfgs, nsteps, stride)
