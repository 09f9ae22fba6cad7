"""torch.autograd.Function over the discrete adjoint of the RK2 step.

    UN = rollout(U0, eta, cfgs, nsteps)

U0 is a float64 tensor (B, nx, ny, 8) of initial states in the reference's
AoS order (ghosts included; the boundary is applied before the first step),
eta a float64 tensor (B, nx, ny) of per-cell resistivities, cfgs B
configurations of the same grid.  The forward runs the HIP library's batched
step with every dt recorded and sqrt(nsteps) checkpoints; backward runs the
hand-written discrete adjoint (mhd_adjoint_backward) with those dt frozen and
returns dL/dU0 and dL/deta.  Compose U0 from primitive fields with torch ops
to get gradients with respect to the initial velocity.  degenerate_tol > 0
differentiates degenerate HLLD star states by their Miyoshi-Kusano limit
(ResistiveMHD2DAdjoint.set_degenerate_tol); 0 differentiates the step
exactly.
"""
import numpy as np
import torch

from .mhd2d import ResistiveMHD2DAdjoint


class MHDRollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, U0, eta, cfgs, nsteps, stride, degenerate_tol):
        if U0.dtype != torch.float64 or eta.dtype != torch.float64:
            raise TypeError("the step is FP64: pass float64 tensors")
        B = U0.shape[0]
        if len(cfgs) != B or eta.shape[0] != B:
            raise ValueError("one configuration and one eta field per run")
        for r, c in enumerate(cfgs):
            if tuple(U0.shape[1:]) != (c.nx, c.ny, 8) or tuple(eta.shape[1:]) != (c.nx, c.ny):
                raise ValueError(f"run {r}: U0 {tuple(U0.shape[1:])} and eta {tuple(eta.shape[1:])} do not match "
                                 f"the grid ({c.nx}, {c.ny})")
        if U0.is_cuda != eta.is_cuda:
            raise ValueError("U0 and eta must be on the same device")
        adj = ResistiveMHD2DAdjoint(cfgs)
        adj.set_degenerate_tol(degenerate_tol)
        if U0.is_cuda:
            # states and resistivities move between torch's memory and the
            # library's on the GPU, with no host round trip
            adj.setEtaField_all_device(eta.detach().contiguous().reshape(-1))
            adj.setU0_all_device(U0.detach().contiguous().reshape(-1))
            adj.forward(int(nsteps), int(stride))
            UN = torch.empty_like(U0)
            adj.getU_all_device(UN.view(-1))
        else:
            etan = eta.detach().contiguous().numpy()
            for r in range(B):
                adj.setEtaField(r, etan[r])
            U0n = U0.detach().contiguous().numpy()
            adj.setU0_all([U0n[r] for r in range(B)])
            adj.forward(int(nsteps), int(stride))
            UN = torch.from_numpy(np.stack(adj.getU_all()))
        ctx.adj = adj
        return UN

    @staticmethod
    def backward(ctx, gUN):
        adj = ctx.adj
        g = gUN.detach().contiguous()
        if g.is_cuda:
            adj.set_seed_all_device(g.reshape(-1))
            adj.backward()
            gU0 = torch.empty_like(g)
            adj.grad_state_all_device(gU0.view(-1))
            geta = torch.empty(g.shape[:-1], dtype=g.dtype, device=g.device)
            adj.grad_eta_all_device(geta.view(-1))
        else:
            gn = g.numpy()
            adj.set_seed_all([gn[r] for r in range(len(adj))])
            adj.backward()
            gU0 = torch.from_numpy(np.stack(adj.grad_state_all()))
            geta = torch.from_numpy(np.stack(adj.grad_eta_all()))
        return gU0, geta, None, None, None, None


def rollout(U0, eta, cfgs, nsteps, stride=0, degenerate_tol=0.0):
    """Final states after nsteps RK2 steps, differentiable in U0 and eta.
    stride: checkpoint stride (0: ceil(sqrt(nsteps)); -1: the largest whose
    storage fits in half of the free device memory, so that the backward
    pass recomputes nothing when the whole rollout fits)."""
    return MHDRollout.apply(U0, eta, cfgs, nsteps, stride, float(degenerate_tol))
