"""Differentiable pieces of the etch step (torch.autograd.Function over hand-written HIP adjoints).

The reference differentiates nothing inside its etch loop (its only gradients are the demo tapes
scripts/inverse_design.py:95-101 and scripts/warp_3d_flux.py:106-113). Here the level-set
advection carries an adjoint: `advect(phi, F, mask, mask_phi, dx, dt, nsub)` runs the same
substeps as ops.advect and back-propagates through them with petch_advect_adjoint, the upwind
choices frozen at each substep's state (the states are kept from the forward pass, nsub x the
grid). Gradients flow to phi and to the speed F; the mask and its pinned values are constants.

    phi1 = adjoint.advect(phi0.requires_grad_(), F.requires_grad_(), mask, mask_phi, dx, dt, 3)
    loss = (phi1 * w).sum(); loss.backward()          # phi0.grad, F.grad

The radiosity solve of mc_flux_3d_radiosity (threed.py:965-985) carries its implicit adjoint:
`radiosity(bare, D, hitface, row_ptr, col, n_ff, beta, n_iter, tol)` returns
m = clip(G, 0, 8) of G = D + Aff diag(1 - clip(bare beta, 0, 1)) G and back-propagates through
the transposed solve (petch_radiosity_adjoint) to bare and to the escape fractions D. The form
factors (hitface, the CSR) are constants: they come from discrete first hits.
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


class _Radiosity(torch.autograd.Function):
    @staticmethod
    def forward(ctx, bare, D, hitface, row_ptr, col, n_ff, beta, n_iter, tol):
        nf = bare.numel()
        b = bare.detach().contiguous()
        m, G, _ = ops.radiosity_solve(row_ptr, col, D.detach(), nf, n_ff, b, beta, n_iter, tol)
        ctx.save_for_backward(b, G, hitface)
        ctx.args = (nf, int(n_ff), float(beta), int(n_iter), float(tol))
        return m

    @staticmethod
    def backward(ctx, gm):
        b, G, hitface = ctx.saved_tensors
        nf, n_ff, beta, n_iter, tol = ctx.args
        gD, gbare, _ = ops.radiosity_adjoint(hitface, nf, n_ff, G, b, beta,
                                             gm.contiguous().float(), n_iter, tol)
        return gbare, gD, None, None, None, None, None, None, None


def radiosity(bare, D, hitface, row_ptr, col, n_ff, beta, n_iter=0, tol=1e-10):
    """m = clip(G, 0, 8) of the radiosity solve (ops.radiosity_solve on the CSR that
    ops.radiosity_assemble built from hitface), differentiable in bare and D. n_iter <= 0 solves
    to max|dG| <= tol max|G|, where the implicit adjoint is exact."""
    return _Radiosity.apply(bare, D, hitface, row_ptr, col, n_ff, beta, n_iter, tol)
