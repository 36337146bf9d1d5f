"""torch.autograd bridge over the adjoint entries.

`thrust_loss(sim, params, nsteps, ...)` is the loss L(p) of an N-step run
(the time-averaged flow-direction Maxwell-stress flux through a sphere around
the dipole) as a differentiable torch function of p = (Mx, My, Mz, V), the
dipole moment vector and the inflow speed.  Its forward is pic_adj_forward
(the context's own fused step, checkpointed) and its backward is
pic_adj_backward (recompute each segment, reverse it in FP64), so it composes
with any torch expression of the loss and of the parameters.

The context keeps the checkpoints of its last forward only: each
`thrust_loss` must be back-propagated before the next one on the same
Simulation, and at most once.  This module is the only one of the package
that imports torch.
"""
import numpy as np
import torch


class _AdjointLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, sim, nsteps, ckpt_every, radius, nquad):
        p = params.detach().to("cpu", torch.float64).numpy().reshape(4)
        loss = sim.adj_forward(p, nsteps, ckpt_every, radius, nquad)
        # sequence number of this forward; the Simulation remembers the one
        # whose checkpoints it holds until that one is back-propagated
        sim._adj_seq = getattr(sim, "_adj_seq", 0) + 1
        ctx.sim = sim
        ctx.token = sim._adj_live = sim._adj_seq
        return params.new_tensor(loss)

    @staticmethod
    def backward(ctx, grad):
        if getattr(ctx.sim, "_adj_live", None) != ctx.token:
            raise RuntimeError(
                "thrust_loss: another forward ran on this Simulation since this "
                "loss was computed; back-propagate each loss before the next")
        ctx.sim._adj_live = None
        g = ctx.sim.adj_backward(float(grad))
        return (torch.as_tensor(np.asarray(g), dtype=grad.dtype, device=grad.device),
                None, None, None, None, None)


def thrust_loss(sim, params, nsteps, ckpt_every=None, radius=3.0, nquad=64):
    """L(p) over `nsteps` steps of `sim` from its current state; params is a
    tensor of 4 values (Mx, My, Mz, V).  ckpt_every defaults to nsteps (one
    segment)."""
    if params.numel() != 4:
        raise ValueError("params must hold 4 values (Mx, My, Mz, V)")
    return _AdjointLoss.apply(params, sim, int(nsteps),
                              int(ckpt_every or nsteps), float(radius), int(nquad))
