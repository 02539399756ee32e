"""torch.autograd surface of one SSA step (glide/torch.py:4-59).

``GlideStep.apply(t, dt, model, level, H_prev, bed, beta, smb) -> (u, v, H)``
runs the native forward solve; its backward runs the native adjoint solve and
returns the gradients with respect to (H_prev, bed, beta, smb).  As in the
reference, the forward solve is warm-started from the velocities the model
currently holds, and H starts from H_prev.
"""
from __future__ import annotations

import torch


class GlideStep(torch.autograd.Function):

    @staticmethod
    def forward(ctx, t, dt, model, level, H_prev, bed, beta, smb):
        ctx.t, ctx.dt, ctx.model, ctx.level = t, dt, model, level
        model.set_top_level(level)
        mg = model.mg
        mg.state.H_prev.set(H_prev.detach(), start_level=level)
        mg.state.H.set(H_prev.detach(), start_level=level)
        mg.geometry.bed.set(bed.detach(), start_level=level)
        mg.sliding.beta.set(beta.detach(), start_level=level)
        mg.forcing.smb.set(smb.detach(), start_level=level)
        model.forward(t, dt, update_geometry=False)
        st = mg[level].state
        u, v, H = st.u.data.clone(), st.v.data.clone(), st.H.data.clone()
        mask, phi = st.mask.data.clone(), st.phi.data.clone()
        ctx.save_for_backward(u, v, H, mask, phi, H_prev, bed, beta, smb)
        return u, v, H

    @staticmethod
    def backward(ctx, gu, gv, gH):
        model, level, dt = ctx.model, ctx.level, ctx.dt
        u, v, H, mask, phi, H_prev, bed, beta, smb = ctx.saved_tensors
        mg = model.mg
        mg.state.H_prev.set(H_prev.detach(), start_level=level)
        mg.geometry.bed.set(bed.detach(), start_level=level)
        mg.sliding.beta.set(beta.detach(), start_level=level)
        mg.forcing.smb.set(smb.detach(), start_level=level)
        mg.state.u.set(u, start_level=level)
        mg.state.v.set(v, start_level=level)
        mg.state.H.set(H, start_level=level)
        mg.state.phi.set(phi, start_level=level)
        mg.state.mask.set(mask, start_level=level)
        zeros = lambda t: torch.zeros_like(t)
        model.backward(ctx.t, dt, dJdu=gu if gu is not None else zeros(u), dJdv=gv if gv is not None else zeros(v),
                       dJdH=gH if gH is not None else zeros(H))
        g = mg[level]
        return (None, None, None, None, g.state.H_prev.grad.clone(), g.geometry.bed.grad.clone(),
                g.sliding.beta.grad.clone(), g.forcing.smb.grad.clone())
