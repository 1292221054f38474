"""torch.autograd bridge for the adjoint xi sweep (qss_sweep_vjp).

    import torch                      # before tertia_amd: one HIP runtime
    from tertia_amd.autograd import SweepEx, witness_loss
    bins = torch.tensor(p.bins, dtype=torch.float64, requires_grad=True)
    ex = SweepEx.apply(bins, mesh, nrun, 32)          # (nrun, nz, ny) on cuda
    loss = witness_loss(ex, p, 6.0, 6.5, 0.2)
    loss.backward()                                   # bins.grad via the adjoint

The forward runs the checkpointed sweep on the library's stream (`every`
layers between checkpoints; 0: every layer's tape recorded, so that the
backward recomputes nothing) and copies the Ex history (the longitudinal
field) into a torch tensor; the backward
hands dLoss/dEx (a device tensor) to the hand-written adjoint and returns
dLoss/dbins.  Both directions synchronize the device around the library
calls, which run on the library's own stream.  One forward per mesh may be
awaiting its backward: the reverse sweep recomputes from the mesh's latest
checkpoints, so a second SweepEx.apply on the same mesh before the first
backward makes that backward raise instead of returning the wrong gradient
(the reverse reads the mesh's latest checkpoints or tape).
"""
from __future__ import annotations

import torch

from .loss import window_mask, witness_loss_np  # noqa: F401


class SweepEx(torch.autograd.Function):
    @staticmethod
    def forward(ctx, bins: torch.Tensor, mesh, nrun: int, every: int = 32):
        mesh.set_bins(bins.detach().to("cpu", torch.float64).numpy())
        mesh.sweep_checkpointed(int(nrun), int(every))
        out = torch.empty((int(nrun), mesh.nz, mesh.ny), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        mesh.copy_history("Ex", 0, int(nrun), out.data_ptr())
        ctx.mesh = mesh
        ctx.generation = mesh.checkpoint_generation
        ctx.bins_device = bins.device
        return out

    @staticmethod
    def backward(ctx, grad_ex: torch.Tensor):
        if ctx.mesh.checkpoint_generation != ctx.generation:
            raise RuntimeError("SweepEx.backward: the mesh ran another checkpointed sweep or "
                               "changed its profile bins after this forward; its adjoint "
                               "would belong to that sweep (use one mesh per graph node)")
        g = grad_ex.detach().to(torch.float64).contiguous()
        torch.cuda.synchronize()
        gb = ctx.mesh.sweep_vjp(device_ptr=g.data_ptr())
        return torch.from_numpy(gb).to(ctx.bins_device), None, None, None


def witness_loss(ex: torch.Tensor, p, xi0: float = 6.0, xi1: float = 6.5,
                 rmax: float = 0.2, var_weight: float = 1e-3) -> torch.Tensor:
    """-mean(Ex) + var_weight * var(Ex) over the witness window."""
    m = torch.from_numpy(window_mask(p, ex.shape[0], xi0, xi1, rmax)).to(ex.device)
    v = ex[m]
    mean = v.mean()
    return -mean + var_weight * ((v - mean) ** 2).mean()
