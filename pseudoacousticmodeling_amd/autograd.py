"""Differentiable forward modelling: ``torch.autograd.Function`` over the C ABI.

``seismograms(vp, engine)`` returns the gathers pam_forward computes for the
interior model ``vp`` ([nz][nx], float32, on the engine's device) and, under
autograd, back-propagates through pam_forward_adjoint, the exact discrete
adjoint of the same time stepping (acoustic physics, Cerjan or CPML; the CPML
profiles' dependence on vp included).  Any loss built from the gathers with
torch operations can be differentiated with respect to vp:

    vp = torch.tensor(v0, device="cuda", requires_grad=True)
    seis = seismograms(vp, eng)
    loss = 0.5 * ((seis - dobs) ** 2).sum()
    loss.backward()          # vp.grad = dloss/dvp

The reference's FWI gradient (src/Migration.py:1039-1092, pam_gradient) is a
preconditioned, illumination-normalised approximation of the gradient with
respect to m = 1/vp^2; this one is the gradient of the discrete map itself.
"""
from __future__ import annotations

import torch

from . import _lib as L


class ForwardModeling(torch.autograd.Function):
    """seis[nshots][nt_data][nrec] = F(vp); backward: vp.grad += F'(vp)^T grad."""

    @staticmethod
    def forward(ctx, vp, engine, shot0, nshots):
        if engine.physics != L.PAM_ACOUSTIC:
            raise NotImplementedError("the exact adjoint covers acoustic physics")
        engine.set_model(vp.detach())
        seis = engine.forward(shot0, nshots)
        ctx.engine = engine
        ctx.shot0, ctx.nshots = shot0, nshots
        ctx.save_for_backward(vp.detach())
        return seis

    @staticmethod
    def backward(ctx, grad_seis):
        (vp,) = ctx.saved_tensors
        eng = ctx.engine
        # the context may have run another model since the forward
        eng.set_model(vp)
        g = eng.forward_adjoint(grad_seis.detach().to(torch.float32).contiguous(), ctx.shot0, ctx.nshots,
                                layout=L.PAM_INTERIOR)
        return g.to(vp.dtype), None, None, None


def seismograms(vp, engine, shot0=0, nshots=None):
    """Gathers of shots shot0 .. shot0+nshots-1 for the interior model vp."""
    nshots = engine.nshot - shot0 if nshots is None else nshots
    return ForwardModeling.apply(vp, engine, int(shot0), int(nshots))


def misfit(vp, engine, dobs, shot0=0, nshots=None):
    """J = 1/2 sum (seis - dobs)^2 as a differentiable scalar (float32 on the
    device; the sum itself is reduced by torch)."""
    seis = seismograms(vp, engine, shot0, nshots)
    r = seis - dobs
    return 0.5 * torch.sum(r * r)
