"""Differentiable pieces of the etch step (torch.autograd.Function over hand-written HIP adjoints).

The reference differentiates nothing inside its etch loop (its only gradients are the demo tapes
scripts/inverse_design.py:95-101 and scripts/warp_3d_flux.py:106-113). Here the level-set
advection carries an adjoint: `advect(phi, F, mask, mask_phi, dx, dt, nsub)` runs the same
substeps as ops.advect and back-propagates through them with petch_advect_adjoint, the upwind
choices frozen at each substep's state (the states are kept from the forward pass, nsub x the
grid). Gradients flow to phi and to the speed F; the mask and its pinned values are constants.

    phi1 = adjoint.advect(phi0.requires_grad_(), F.requires_grad_(), mask, mask_phi, dx, dt, 3)
    loss = (phi1 * w).sum(); loss.backward()          # phi0.grad, F.grad
"""
import torch

from . import ops


class _Advect(torch.autograd.Function):
    @staticmethod
    def forward(ctx, phi, F, mask, mask_phi, dx, dt, nsub):
        states = [phi.detach().contiguous()]
        cur = states[0]
        for _ in range(int(nsub)):
            nxt = cur.clone()
            ops.advect(nxt, F.detach(), mask, mask_phi, dx, dt, 1)
            states.append(nxt)
            cur = nxt
        ctx.save_for_backward(*states[:-1], F.detach(), mask)
        ctx.dx, ctx.dt = float(dx), float(dt)
        return cur.clone()

    @staticmethod
    def backward(ctx, gout):
        *states, F, mask = ctx.saved_tensors
        gF = torch.zeros_like(F)
        g = gout.contiguous()
        for s in reversed(states):
            g = ops.advect_adjoint(s, F, mask, g, ctx.dx, ctx.dt, gF)
        return g, gF, None, None, None, None, None


def advect(phi, F, mask, mask_phi, dx, dt, nsub):
    """nsub upwind substeps of size dt with the mask re-pin (ops.advect), differentiable in phi
    and F."""
    return _Advect.apply(phi, F, mask, mask_phi, dx, dt, nsub)
