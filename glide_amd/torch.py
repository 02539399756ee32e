"""torch.autograd surfaces: one SSA step (glide/torch.py:4-59) and one
thermal step.

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


class ThermalStep(torch.autograd.Function):
    """torch.autograd surface of one thermal step (no counterpart in the
    reference, whose enthalpy solve has no adjoint).

    ``ThermalStep.apply(thermal, dt, E_prev, Q_geo, E_surface) -> (E, B)`` sets
    the previous enthalpy (also the initial iterate), the geothermal flux and
    the surface enthalpy of ``thermal`` (a ``ThermalModel`` after its momentum
    step and ``pre_momentum`` snapshot of H), runs ``thermal.step(dt)`` and
    returns the new enthalpy and rate factor (head units; zeros when
    ``update_rheology`` is off).  The backward solves the adjoint of the
    converged solve at the returned state, so it is exact when the forward
    sweeps converge (tighten ``thermal.ops.smoother_config`` tolerances and
    ``thermal.n_smooth`` for gradient work); the thermal model must not step
    again between the two passes."""

    @staticmethod
    def forward(ctx, thermal, dt, E_prev, Q_geo, E_surface):
        ops = thermal.ops
        ops.enthalpy_state.E_prev.copy_(E_prev.detach())
        ops.enthalpy_state.E.copy_(E_prev.detach())
        ops.enthalpy_forcing.Q_geo.copy_(Q_geo.detach())
        ops.enthalpy_forcing.E_surface.copy_(E_surface.detach())
        thermal.step(dt)
        E = ops.enthalpy_state.E.clone()
        B = ops.grid.rheology.B.data.clone() if thermal.update_rheology else torch.zeros_like(Q_geo)
        ctx.thermal, ctx.dt = thermal, dt
        return E, B

    @staticmethod
    def backward(ctx, gE, gB):
        thermal = ctx.thermal
        ops = thermal.ops
        if gE is None:
            gE = torch.zeros_like(ops.enthalpy_state.E)
        use_B = gB is not None and thermal.update_rheology
        g = ops.adjoint_solve(ctx.dt, gE.contiguous(), gB.contiguous() if use_B else None, B_scale=thermal.B_scale)
        return None, None, g["E_prev"], g["Q_geo"], g["E_surface"]
